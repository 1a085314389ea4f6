// Library-internal entry points between the C++ learned pipeline and the
// engine and network objects it drives.  Not part of the public C ABI: the
// pipeline uses them to hand the clouds that the refine engine has uploaded
// to the network as a copy on the device, instead of a second upload from
// host memory.
#ifndef RTR_INTERNAL_H
#define RTR_INTERNAL_H

#include <stddef.h>

#include "../../include/rtr.h"

#ifdef __cplusplus
extern "C" {
#endif

// The engine's resident clouds as packed xyz on its device (12 B per point):
// every source in order, and every target in order.
int rtr_batch_device_clouds(const rtr_batch* b, const float** src_xyz, const float** tgt_xyz);

// Clouds already on the network's device, packed xyz: the sources of pairs
// 0..n_pairs-1 one after another from src_xyz, the targets likewise from
// tgt_xyz.  Copied on the device; the same resident clouds as
// rtr_prop_set_clouds with the same points.
int rtr_prop_set_device_clouds(rtr_prop* h, int n_pairs, const float* src_xyz, const size_t* ns,
                               const float* tgt_xyz, const size_t* nt);

#ifdef __cplusplus
}
#endif

#endif
