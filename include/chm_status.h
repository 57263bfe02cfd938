/*
 * Status codes shared by every C entry point of charming_amd.
 * The reference reports failures with printf + exit(-1) on the host
 * (src/util/util.h:6-12, src/charming.cu:95-129) and printf + assert on the
 * device (src/comm/get/get.cu:738-742); here every failure is a negative
 * return code plus a message string.
 */
#ifndef CHM_STATUS_H_
#define CHM_STATUS_H_

#define CHM_OK 0
#define CHM_E_INVALID (-1)     /* bad argument or shape */
#define CHM_E_HIP (-2)         /* a HIP runtime call failed */
#define CHM_E_NOMEM (-3)       /* device or host allocation failed */
#define CHM_E_UNSUPPORTED (-4) /* feature outside this build */
#define CHM_E_DEVICE (-5)      /* the device runtime raised its error word */
#define CHM_E_TIMEOUT (-6)     /* host watchdog expired; the kernel was told to stop */
#define CHM_E_STATE (-7)       /* call out of order (e.g. run before init) */
#define CHM_E_NOGPU (-8)       /* no usable gfx950 device */
#define CHM_E_BUSY (-9)        /* admission refused for now: budget or slots in use */

#endif /* CHM_STATUS_H_ */
