// rt_nn_abi.h -- internal glue between the denoiser ABI and the renderer.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rt/rt_hip.h"
#include "rt_nn.h"

// sets the thread-local message returned by rt_last_error()
void rt_set_error(const char* msg);

// inference of `nn` over P pixels of device frame buffers on `st` (no
// synchronisation); the caller keeps calls on one network from overlapping
hipError_t rt_nn_infer_on_stream(rt_nn* nn, const rtk::NnImages& images, int P,
    const float4* beauty, const float4* normal, const float4* albedo, const float* depth,
    float4* out, hipStream_t st);
int rt_nn_device(const rt_nn* nn);
// reference counting: a renderer keeps the network it uses alive
void rt_nn_retain(rt_nn* nn);
void rt_nn_release(rt_nn* nn);
// allocates the network's inference scratch for these images (before a graph capture)
hipError_t rt_nn_reserve(rt_nn* nn, const rtk::NnImages& images);
