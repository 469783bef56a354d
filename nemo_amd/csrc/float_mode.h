/*
 * Launch interface of the float-weight mode kernels (float_mode.hip): plain
 * C ABI, device pointers, sizes and a stream; each returns the launch status.
 */
#ifndef NEMO_AMD_FLOAT_MODE_H
#define NEMO_AMD_FLOAT_MODE_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One forward step t of the float-weight network (all N neurons). */
typedef struct nemo_float_step_args {
	unsigned n;                /* neurons */
	unsigned t;                /* step */
	const uint32_t* rcm_ptr;   /* n+1 offsets of each target's incoming synapses */
	const uint32_t* rcm_sd;    /* source | (delay-1) << 26 */
	const float* wr;           /* weights in reverse-matrix order */
	const float* pa;           /* Izhikevich a, b, c, d per neuron */
	const float* pb;
	const float* pc;
	const float* pd;
	const float* x;            /* [T][n] external current */
	float* S;                  /* [T][n] spikes (written at t when store) */
	float* V;                  /* [T][n] membrane potential after the step (store) */
	double* v;                 /* [n] state, updated */
	double* u;
	int store;                 /* write S[t], V[t] and the bitmap of step t */
	double* ckpt;              /* NULL, or [ceil(T/K)][2][n]: (v, u) before every K-th step */
	unsigned K;
	uint32_t* bits;            /* [T][pitch] spike bitmaps (read for t-d, written at t when store) */
	unsigned pitch;            /* words per step, nemo_float_bitmap_pitch(n) or more, a multiple of 4 */
	unsigned dmax;             /* longest delay */
	int stage;                 /* stage the last dmax bitmaps in LDS; needs nemo_float_stage_bytes
	                              of at most NEMO_FLOAT_STAGE_MAX */
} nemo_float_step_args;

/* The static workspace of the allocation-free forward: a file-scope device
 * array of NEMO_FLOAT_SCRATCH_BYTES, one per device (each device holds it
 * until the process ends).  Its address on the current device, looked up
 * once per device (with hipGetSymbolAddress, the only valid device address
 * of the array in host code); devices numbered below
 * NEMO_FLOAT_SCRATCH_DEVICES. */
#define NEMO_FLOAT_SCRATCH_BYTES (64u << 20)
#define NEMO_FLOAT_SCRATCH_DEVICES 64
hipError_t nemo_float_scratch(void** ptr, size_t* bytes);

/* Largest LDS stage of spike bitmaps a forward block takes. */
#define NEMO_FLOAT_STAGE_MAX 49152
unsigned nemo_float_bitmap_pitch(unsigned n);
size_t nemo_float_stage_bytes(unsigned n, unsigned dmax);

/* Replay of forward steps t0..t1-1 from the state before t0 and the
 * recorded spikes (as firing histories): each step's entry state and input,
 * equal to the forward's. */
typedef struct nemo_float_replay_args {
	unsigned n;
	unsigned t0, t1;
	unsigned dmax;             /* longest delay, at most 64 */
	const uint32_t* rcm_ptr;
	const uint32_t* rcm_sd;
	const float* wr;
	const float* pa;
	const float* pb;
	const float* pc;
	const float* pd;
	const float* x;            /* [T][n] */
	const uint64_t* hist;      /* [hist_rows][n] from nemo_launch_pack_history */
	unsigned hist_rows;
	const double* v0;          /* [n] state before t0 */
	const double* u0;
	double* seg_v;             /* [t1 - t0][n] entry state and input of each step */
	double* seg_u;
	double* seg_I;
} nemo_float_replay_args;

hipError_t nemo_launch_float_replay(hipStream_t, const nemo_float_replay_args*);
/* rows of the firing histories of a T-step run */
unsigned nemo_float_history_rows(unsigned T);
/* hist[r][j], bit k: neuron j fired at step 64 (r - 1) + k (S [T][n]); row 0
 * and steps from T on are zero */
hipError_t nemo_launch_pack_history(hipStream_t, uint64_t* hist, unsigned rows, const float* S, unsigned n,
		unsigned T);
/* One backward step t (in decreasing t) of the adjoint. */
typedef struct nemo_float_back_args {
	unsigned n;
	unsigned t;
	unsigned T;                /* steps of the forward run */
	unsigned ring;             /* slots of LI (max delay + 1) */
	const uint32_t* fcm_ptr;   /* n+1 offsets of each source's synapses */
	const uint32_t* fcm_td;    /* target | (delay-1) << 26 */
	const float* wf;           /* weights in forward-matrix order */
	const float* pa;
	const float* pb;
	const float* pc;
	const float* pd;
	const float* S;            /* [T][n] forward spikes */
	const float* gS;           /* [T][n] dL/dS, or NULL */
	const float* gV;           /* [T][n] dL/dV, or NULL */
	unsigned t0;               /* first step of the recorded segment */
	const double* seg_v;       /* recorded entry state and input of the segment */
	const double* seg_u;
	const double* seg_I;
	double* LV;                /* [n] dL/dv of the state after step t; becomes the one before */
	double* LU;
	double* LI;                /* [ring][n] dL/dI of recent steps */
	float* gw;                 /* [synapses] dL/dw in forward-matrix order, accumulated */
	float* ga;                 /* [n] accumulated */
	float* gb;
	float* gd;
	float* gx;                 /* [T][n] dL/dx, or NULL */
} nemo_float_back_args;

hipError_t nemo_launch_float_forward_step(hipStream_t, const nemo_float_step_args*);
hipError_t nemo_launch_float_backward_step(hipStream_t, const nemo_float_back_args*);
/* out[k] = in[idx[k]] */
hipError_t nemo_launch_gather(hipStream_t, float* out, const float* in, const uint32_t* idx, size_t n);
/* out[idx[k]] = in[k] (idx a permutation) */
hipError_t nemo_launch_scatter(hipStream_t, float* out, const float* in, const uint32_t* idx, size_t n);
hipError_t nemo_launch_fill(hipStream_t, float* p, float v, size_t n);
hipError_t nemo_launch_fill_double(hipStream_t, double* p, double v, size_t n);
/* out[k] = float(in[k]) */
hipError_t nemo_launch_narrow(hipStream_t, float* out, const double* in, size_t n);
/* out[k] = double(in[k]) */
hipError_t nemo_launch_widen(hipStream_t, double* out, const float* in, size_t n);

#ifdef __cplusplus
}
#endif

#endif
