"""CPU model of the runtime's inbox protocol under random interleavings.

Restates, step by step, what include/charming_detail/device_rt.h does for one
destination PE: producers reserve with a fetch-add on the tail, wait while
the reservation would overrun the head (giving up after a spin limit, or at
once on a send to self), write the message and publish the tag of its first
64-byte granule (tag = granule number + 1, never reused); the consumer polls
the tag of its read position, runs the entry method through the refnum gate
or copies the message into the deferral arena (compacting when the arena is
full, failing when it stays full), releases the ring space by advancing the
head, and re-dispatches revived messages (set_refnum) before new ones in
arrival order.  Every step is a separate scheduling point, so producers and
the consumer interleave arbitrarily.  The checks: each message is delivered
exactly once, in refnum order, with its own payload; the ring drains; the
error paths fire with the device's error codes.
"""
import random

import pytest

GRANULE = 64
ERR_SELF_RING_FULL, ERR_DEFERRED_FULL, ERR_SPIN_LIMIT = 2, 3, 8


def alloc_size(payload):
    s = 16 + 16 + payload  # envelope + regular_msg + payload
    return (s + 15) // 16 * 16


class Error(Exception):
    def __init__(self, code):
        super().__init__(code)
        self.code = code


class Inbox:
    """One PE's ring, publish words and deferral arena (device_rt.h)."""

    def __init__(self, ring_bytes, deferred_bytes, deferred_max=1024):
        self.R = ring_bytes
        self.NG = ring_bytes // GRANULE
        self.tail = 0
        self.head = 0
        self.ready = [0] * self.NG
        self.mem = {}            # ring position -> message (logical address)
        self.read_pos = 0
        self.arena_bytes = deferred_bytes
        self.arena_used = 0
        self.deferred = []       # [offset, bytes, msg, state, seq]
        self.deferred_max = deferred_max
        self.seq = 0
        self.error = 0

    def raise_error(self, code):
        if not self.error:
            self.error = code
        raise Error(code)


def producer(box, me, dst_me, msgs, spin_limit):
    """Generator: one PE sending msgs to the box (reserve/write/publish)."""
    for m in msgs:
        slot = (alloc_size(m["bytes"]) + GRANULE - 1) // GRANULE * GRANULE
        pos = box.tail                        # fetch_add on the tail
        box.tail += slot
        yield
        spins = 0
        while pos + slot - box.head > box.R:  # would overrun the head
            if dst_me:
                box.raise_error(ERR_SELF_RING_FULL)
            spins += 1
            if spin_limit and spins > spin_limit:
                box.raise_error(ERR_SPIN_LIMIT)
            yield
        box.mem[pos] = dict(m, slot=slot)     # envelope, header, payload
        yield
        g = pos // GRANULE                    # release, then the tag
        box.ready[g % box.NG] = g + 1
        yield


def consumer(box, chare, deliver, n_total):
    """Generator: the destination PE's scheduler loop."""
    done = 0
    while done < n_total:
        revived = [d for d in box.deferred if d[3] == 2]
        if revived:
            d = min(revived, key=lambda e: e[4])
            m = d[2]
            if chare.gate(m["refnum"]):
                box.deferred.remove(d)
                if not box.deferred:
                    box.arena_used = 0
                deliver(m, chare)
                done += 1
            else:
                d[3] = 1
            yield
            continue
        g = box.read_pos // GRANULE
        if box.ready[g % box.NG] != g + 1:    # not published yet
            yield
            continue
        m = box.mem.pop(box.read_pos)         # acquire: the message is whole
        if chare.gate(m["refnum"]):
            deliver(m, chare)
            done += 1
        else:
            defer(box, m)
        yield
        box.read_pos += m["slot"]             # release the ring space
        box.head = box.read_pos
        yield


def defer(box, m):
    need = (alloc_size(m["bytes"]) + 15) // 16 * 16
    if box.arena_used + need > box.arena_bytes:   # compact live entries
        off = 0
        for d in sorted(box.deferred, key=lambda e: e[0]):
            d[0] = off
            off += d[1]
        box.arena_used = off
    if (len(box.deferred) >= box.deferred_max or
            box.arena_used + need > box.arena_bytes):
        box.raise_error(ERR_DEFERRED_FULL)
    box.deferred.append([box.arena_used, need, m, 1, box.seq])
    box.seq += 1
    box.arena_used += need


class Chare:
    def __init__(self, box):
        self.box = box
        self.refnum = 0
        self.got = []

    def gate(self, ref):
        return ref == -1 or ref == self.refnum

    def set_refnum(self, v):
        self.refnum = v
        for d in self.box.deferred:          # revive_mismatches
            if d[3] == 1 and d[2]["refnum"] == v:
                d[3] = 2


def deliver_in_order(m, chare):
    assert m["payload"] == ("p", m["src"], m["k"], m["bytes"])
    chare.got.append(m["refnum"])
    chare.set_refnum(chare.refnum + 1)


def run(gens, rng, max_steps=2_000_000):
    live = list(gens)
    steps = 0
    while live:
        g = rng.choice(live)
        try:
            next(g)
        except StopIteration:
            live.remove(g)
        steps += 1
        assert steps < max_steps, "no progress"
    return steps


def make_msgs(n_prod, per, rng, sizes):
    """refnums 0..N-1 spread over the producers in a random order, so most
    messages arrive ahead of the consumer's refnum and are deferred."""
    refs = list(range(n_prod * per))
    rng.shuffle(refs)
    out = []
    for p in range(n_prod):
        out.append([{"src": p, "k": k, "refnum": refs[p * per + k],
                     "bytes": rng.choice(sizes),
                     "payload": ("p", p, k, None)} for k in range(per)])
        for m in out[-1]:
            m["payload"] = ("p", p, m["k"], m["bytes"])
    return out


@pytest.mark.parametrize("seed", range(12))
def test_exactly_once_in_refnum_order(seed):
    rng = random.Random(seed)
    box = Inbox(ring_bytes=4096, deferred_bytes=1 << 20)
    chare = Chare(box)
    n_prod, per = 3, 20
    msgs = make_msgs(n_prod, per, rng, [0, 8, 40, 100, 500, 1000])
    gens = [producer(box, p, False, msgs[p], 0) for p in range(n_prod)]
    gens.append(consumer(box, chare, deliver_in_order, n_prod * per))
    run(gens, rng)
    assert chare.got == list(range(n_prod * per))
    assert box.head == box.tail == box.read_pos   # the ring drained
    assert not box.deferred and box.error == 0


@pytest.mark.parametrize("seed", range(6))
def test_wraps_the_ring_many_times_with_unique_tags(seed):
    # a 1 KiB ring and 200 messages: positions wrap about 30 times; a stale tag of
    # an earlier lap must never look published
    rng = random.Random(100 + seed)
    box = Inbox(ring_bytes=1024, deferred_bytes=1 << 20)
    chare = Chare(box)
    msgs = [[{"src": 0, "k": k, "refnum": k, "bytes": rng.choice([8, 100, 200]),
              "payload": None} for k in range(200)]]
    for m in msgs[0]:
        m["payload"] = ("p", 0, m["k"], m["bytes"])
    run([producer(box, 0, False, msgs[0], 0),
         consumer(box, chare, deliver_in_order, 200)], rng)
    assert chare.got == list(range(200))
    assert box.tail > 20 * box.R


def test_full_ring_hits_the_spin_limit():
    # the consumer never runs (its PE is held): the sender gives up
    box = Inbox(ring_bytes=1024, deferred_bytes=1 << 20)
    msgs = [{"src": 0, "k": k, "refnum": k, "bytes": 200, "payload": None}
            for k in range(10)]
    with pytest.raises(Error) as e:
        run([producer(box, 0, False, msgs, spin_limit=50)], random.Random(1))
    assert e.value.code == ERR_SPIN_LIMIT and box.error == ERR_SPIN_LIMIT


def test_full_ring_on_a_send_to_self_fails_at_once():
    box = Inbox(ring_bytes=1024, deferred_bytes=1 << 20)
    msgs = [{"src": 0, "k": k, "refnum": k, "bytes": 200, "payload": None}
            for k in range(10)]
    with pytest.raises(Error) as e:
        run([producer(box, 0, True, msgs, spin_limit=0)], random.Random(2))
    assert e.value.code == ERR_SELF_RING_FULL


def test_backpressure_without_a_limit_waits_and_completes():
    # a slow consumer and no spin limit: producers wait for ring space
    rng = random.Random(3)
    box = Inbox(ring_bytes=512, deferred_bytes=1 << 20)
    chare = Chare(box)
    msgs = [[{"src": p, "k": k, "refnum": -1, "bytes": 300, "payload": None}
             for k in range(30)] for p in range(2)]
    got = []

    def take(m, c):
        got.append((m["src"], m["k"]))
    run([producer(box, p, False, msgs[p], 0) for p in range(2)] +
        [consumer(box, chare, take, 60)], rng)
    for p in range(2):   # FIFO per producer (reservation order)
        assert [k for s, k in got if s == p] == list(range(30))


def test_deferred_store_overflow():
    # refnums the chare never takes fill a small arena
    box = Inbox(ring_bytes=1 << 16, deferred_bytes=4096)
    chare = Chare(box)
    chare.refnum = -5
    msgs = [{"src": 0, "k": k, "refnum": 1000 + k, "bytes": 1000,
             "payload": None} for k in range(10)]
    with pytest.raises(Error) as e:
        run([producer(box, 0, False, msgs, 0),
             consumer(box, chare, deliver_in_order, 10)], random.Random(4))
    assert e.value.code == ERR_DEFERRED_FULL
    assert len(box.deferred) == 4096 // alloc_size(1000)


def test_compaction_reuses_the_arena():
    # deferred messages revived out of arrival order leave holes; compaction
    # packs the live ones so later deferrals fit
    rng = random.Random(5)
    box = Inbox(ring_bytes=1 << 16, deferred_bytes=3 * alloc_size(500))
    chare = Chare(box)
    order = [2, 1, 0, 5, 4, 3, 8, 7, 6]
    msgs = [[{"src": 0, "k": k, "refnum": r, "bytes": 500, "payload": None}
             for k, r in enumerate(order)]]
    for m in msgs[0]:
        m["payload"] = ("p", 0, m["k"], 500)
    run([producer(box, 0, False, msgs[0], 0),
         consumer(box, chare, deliver_in_order, 9)], rng)
    assert chare.got == list(range(9)) and box.error == 0


def test_planted_lap_free_tag_is_caught():
    # plant a mistake in the model: a consumer that takes any non-zero
    # publish word as "published" (no lap in the tag).  After the ring
    # wraps, a word left by the previous lap makes it read a message that is
    # not written yet, which some interleaving exposes.
    def weak_consumer(box, chare, deliver, n_total):
        done = 0
        while done < n_total:
            g = box.read_pos // GRANULE
            if box.ready[g % box.NG] == 0:
                yield
                continue
            m = box.mem.pop(box.read_pos, None)
            if m is None:
                raise Error(-1)   # read a slot nobody has written yet
            deliver(m, chare)
            done += 1
            yield
            box.read_pos += m["slot"]
            box.head = box.read_pos
            yield
    caught = 0
    for seed in range(20):
        rng = random.Random(200 + seed)
        box = Inbox(ring_bytes=512, deferred_bytes=1 << 20)
        chare = Chare(box)
        msgs = [{"src": 0, "k": k, "refnum": k, "bytes": 100,
                 "payload": ("p", 0, k, 100)} for k in range(60)]
        try:
            run([producer(box, 0, False, msgs, 0),
                 weak_consumer(box, chare, deliver_in_order, 60)], rng)
        except (Error, AssertionError):
            caught += 1
    assert caught > 0
