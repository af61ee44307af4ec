/*
 * msj_params.h -- the reference's params.cuh knobs on the self-join path,
 * under the same names and with the same defaults.  Compile-time in the
 * reference; here they are the defaults of msj_params (runtime values).
 *
 * Reference: include/params.cuh:20-86 (bwd29/self-join-MiSTIC).
 */
#ifndef MSJ_PARAMS_H
#define MSJ_PARAMS_H

/* Node-net construction (nodes.cu) */
#define MAXRP 6                      /* params.cuh:50  layers of the node net          */
#define RPPERLAYER 32                /* params.cuh:53  reference points tried per layer */
#define NUMIDIM 6                    /* params.cuh:61  axis partitions tried per layer  */
#define SAMPLE_PER 0.01              /* params.cuh:69  RP scoring sample: first 1% rows */
#define MIN_CALCS_PER_NODE 100000ull /* params.cuh:38  stop splitting below this       */
#define USE_MIN_CALC 1               /* params.cuh:39                                  */

/* Distance kernel (kernel.cu:833-865) */
#define ILP 4                        /* params.cuh:51  early-exit chunk width (fixed)  */

/* Launcher (launcher.cu:3203-3249); kept for the stderr record's batch count. */
#define KERNEL_BLOCKS 1024           /* params.cuh:65  (KB in the reference Makefile)  */
#define BLOCK_SIZE 256               /* params.cuh:64  (BS in the reference Makefile)  */
#define CALC_PER_BATCH 115ull        /* params.cuh:72                                  */
#define CMP 14                       /* params.cuh:49                                  */

/* Device and output buffers */
#define CUDA_DEVICE 0                /* params.cuh:40 is 1 on the reference's node; 0 here */
#define ERRORPRINT 1                 /* params.cuh:31  whitespace stderr record        */

#endif
