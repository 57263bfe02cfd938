"""Full-size parity through committed golden fixtures: per-chare interior
hashes and the exact last-sweep residual written by the CPU oracle's
message-driven runner (tests/golden/make_fixtures.py)."""
import glob
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "jacobi_*.json")))


def split(length, n, r):
    b = length // n
    return r * b, (length - b * (n - 1) if r == n - 1 else b)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
@pytest.mark.parametrize("form", ["persistent", "plain"])
def test_fixture(path, form):
    import charming_amd as cam
    d = json.load(open(path))
    W, H, (cr, cc), iters, seed = d["W"], d["H"], d["chares"], d["iters"], d["seed"]
    print(f"fixture {os.path.basename(path)} form {form}", flush=True)
    kw = {"chares": (cr, cc)} if form == "persistent" else {}
    with cam.Jacobi2D(W, H, iters, form=form, seed=seed, **kw) as j:
        r = j.solve()
        assert r.residual == int(d["residual_last_sweep"])
        assert r.hash == int(d["grid_hash"])
        bad = []
        for i, hexp in enumerate(d["chare_hashes"]):
            r0, bh = split(H, cr, i // cc)
            c0, bw = split(W, cc, i % cc)
            if j.hash_region(r0 + 1, r0 + 1 + bh, c0 + 1, c0 + 1 + bw) != int(hexp):
                bad.append(i)
        assert not bad, f"chares differing from the oracle: {bad[:20]}"
