// crt_kernels.h - host-side launchers of the gfx950 kernels (kernels.hip). Each returns the
// hipError_t of the launch as an int.

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "crt_layout.h"

// words of one half of the streaming pass control block (two halves alternate)
#define CRT_CTL_WORDS 8
#define CRT_CTL_SURV 0   // survivors written to the front of the other state buffer by the shade stage
#define CRT_CTL_GEN 2    // next path id of the pass to generate
#define CRT_CTL_LIVE 3   // paths in this iteration's state buffer (survivors + new)
#define CRT_CTL_WORK 4   // ray hand-out counter of the persistent closest-hit stage

// SoA path state of one buffer (two buffers alternate between iterations)
struct crt_paths {
    float4* ray_o;   // origin, time
    float4* ray_d;   // direction, state bits (depth | bounce << 8 | camera << 16)
    float4* thr;     // throughput, minstd engine state
    float4* rad;     // radiance gathered so far
    uint2* ids;      // storage pixel, request << 16 | sample
};

extern "C" {
int crt_k_pass_init(hipStream_t s, uint32_t* ctl, uint32_t nwords);
// rng_mode CRT_RNG_PCG / CRT_RNG_MINSTD, or CRT_RNG_DYNAMIC (each path's mode read from its
// request) for launches that mix requests of both kinds
#define CRT_RNG_DYNAMIC 2
int crt_k_regen(hipStream_t s, int rng_mode, uint32_t grid, const crt_dev_request* reqs, int nreq, uint32_t pid0,
                uint32_t total, uint32_t nslots, uint32_t* ctl_in, uint32_t* ctl_out, const uint32_t* open_pass,
                crt_paths P);
int crt_k_extend(hipStream_t s, int count_mode, int width, int two_level, uint32_t grid, const crt_dev_scene* scenes,
                 const crt_dev_request* reqs, int32_t uniform_scene, const uint32_t* queue, const uint32_t* count_ptr, uint32_t count_direct, const float4* ray_o,
                 const float4* ray_d, const uint2* ids, float4* hit, uint32_t* hit_inst, uint32_t* gstack,
                 uint32_t* zero_counter, int mode, uint32_t* work_counter, uint32_t param, int stack_lds,
                 const uint32_t* first_camera, crt_dev_counters* counters);
// first_camera (nullable): queue index where the camera rays of the launch begin (the survivors of
// the last shade stage come first); the lane-refill kernel keeps camera-ray waves whole
// two_level: 0 one-level scenes only, 1 an instanced scene has paths in the launch, 2 and a scene
// with analytic shapes too (the two-level kernels with the shape test; crt_k_extend, crt_k_plain and
// crt_k_trace_rays read it the same way);
// live_min (nullable, grid words): per block, the smallest path id still alive after the stage;
// one_request: every path of the launch belongs to reqs[0]; full: a scene of the launch is instanced
// or has analytic shapes (the unbounded kernel with the shapes' shading)
int crt_k_shade(hipStream_t s, int rng_mode, int full, uint32_t grid, const crt_dev_scene* scenes,
                const crt_dev_request* reqs, const uint32_t* count_in, crt_paths Pin, crt_paths Pout, uint32_t* count_out, const float4* hit,
                const uint32_t* hit_inst, float4* out_rad, uint32_t* live_min, int sort_octant, int one_request,
                crt_dev_counters* counters);
int crt_k_plain(hipStream_t s, int rng_mode, int width, int two_level, uint32_t grid, const crt_dev_scene* scenes,
                const crt_dev_request* req, uint32_t* gstack, crt_dev_counters* counters);
int crt_k_trace_rays(hipStream_t s, int width, int two_level, uint32_t grid, const crt_dev_scene* sc, const float4* rays, uint32_t n, int camera,
                     float4* hit, uint32_t* hit_inst, uint32_t* gstack, crt_dev_counters* counters);
// rays (nreq == 1 only, may be null): the closest-hit queries of the resolved paths are added to it
int crt_k_resolve(hipStream_t s, const crt_dev_request* reqs, int nreq, uint32_t max_pix, const float4* rad,
                  unsigned long long* rays);
int crt_k_tonemap(hipStream_t s, const float4* accum, uint32_t npix, uint32_t spp, uint8_t* rgb);
int crt_k_init_request(hipStream_t s, float4* accum, uint32_t* rng_state, uint32_t W, uint32_t H, uint64_t seed);
int crt_k_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v);
// VALU issue-rate calibration (kind 0: v_fma_f32, 1: v_add_u32, 2: v_mul_f32; 3-7: v_fma_f32 on
// lanes 0-15, every fourth lane, lanes 0-31, 0-62, 0-47 of each wave; 8: v_fma_f32 with EXEC
// cleared; 9: v_cndmask_b32 (one v_cmp per 15), 10: v_max_f32, 11: v_cmp_lt_f32, 12: v_mov_b32,
// 13: v_cndmask_b32_e64 with the mask in an SGPR pair; 14 / 15: v_fma_f32 with EXEC set to lanes
// 0-15 / 0-31 inside the asm):
// blocks x 256 lanes, iters rounds of 16 wave instructions; one float per lane to out
int crt_k_valu_probe(hipStream_t s, int kind, uint32_t blocks, uint32_t iters, float* out);
uint32_t crt_k_stack_capacity(void);
uint32_t crt_k_gstack_entries(void);
}
