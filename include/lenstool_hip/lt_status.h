/*
 * lt_status.h -- status codes and enumerations shared by the C-ABI (lt_capi.h) and its users.
 *
 * The reference has no return codes: its GPU wrappers print and exit(-1) on a CUDA error
 * (src/grid_gradient_GPU.cu:48-51) or ignore errors altogether (src/grid_gradient2_GPU.cu:47).
 * Every lt_* entry point returns one of the codes below instead; lt_last_error() gives the text.
 */
#ifndef LENSTOOL_HIP_LT_STATUS_H
#define LENSTOOL_HIP_LT_STATUS_H

#define LT_OK             0
#define LT_ERR_INVALID    1   /* bad argument: null pointer, non-positive size, non-finite value */
#define LT_ERR_TYPE       2   /* halo type without an implementation for the requested quantity */
#define LT_ERR_HIP        3   /* HIP runtime error (message from hipGetErrorString)             */
#define LT_ERR_NOMEM      4   /* host or device allocation failed                               */
#define LT_ERR_OVERFLOW   5   /* more than LT_MAXIMPERSOURCE images for one source (clamped)    */
#define LT_ERR_CAPACITY   6   /* request larger than the configured HBM budget                  */
#define LT_ERR_CANCELLED  7   /* request cancelled before it completed                          */
#define LT_ERR_NOT_FOUND  8   /* unknown handle / request id                                    */
#define LT_ERR_NO_DEVICE  9   /* no usable GPU                                                  */
#define LT_ERR_BUSY      10   /* result not ready yet (non-blocking poll)                       */

/* arithmetic of the halo sum */
#define LT_PREC_FP64   0      /* plain FP64                                                     */
#define LT_PREC_FP32C  1      /* FP32 with compensated accumulation, FP64 near singular cores  */

/* map kinds (reference runmode keywords ampli / shear / mass / poten / dpl) */
#define LT_MAP_AMPLI    1     /* modes 1..6  (src/grid_map_ampli_GPU.cu:65-90)                  */
#define LT_MAP_SHEAR    2     /* modes 1..4  (src/grid_map_shear_GPU.cu:64-83)                  */
#define LT_MAP_MASS     3     /* modes 1..4  (src/grid_map_mass_GPU.cu:60-76), Kmap Hessian     */
#define LT_MAP_POT      4     /* modes 1..3  (src/grid_map_pot_GPU.cu:64-78)                    */
#define LT_MAP_DPL      5     /* mode 1      (src/grid_map_dpl_GPU.cu:75-146), two planes       */
#define LT_MAP_FIELDS   6     /* kappa, gamma1, gamma2, det A, mu in one pass (five planes)     */
#define LT_MAP_HESSIAN  7     /* raw Hessian planes a, b, c, d                                  */
#define LT_MAP_GRADIENT 8     /* raw gradient planes x, y                                       */

#endif
