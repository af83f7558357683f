// guide.hip -- the path-guiding network on MFMA (include/ymx_guide.h).
//
// Every wave works on tiles of 16 samples with v_mfma_f32_16x16x32_bf16:
// layer l computes H_l^T (64 neurons x 16 samples) = W_l (64 x 64) . H_{l-1}^T.
// A lane (g = lane / 16, s = lane % 16) holds
//   B operand (activations): 8 BF16 of sample s at input positions 8g .. 8g+7,
//   A operand (weights):     8 BF16 of row 16t + s at the same positions,
//   D (accumulator):         4 FP32 of sample s, neurons 16t + 4g .. 16t + 4g + 3.
// The accumulators of tiles 2k and 2k+1 are therefore exactly the B operand
// of K-step k of the next layer once that layer reads its input dimension in
// the permuted order (4g .. 4g+3, 16 + 4g .. 16 + 4g + 3 within each block of
// 32) -- the activations never leave registers between layers.  The encoder
// lays out the first layer's input so that lane group g computes hash-grid
// levels 4g .. 4g+3 of its sample.
//
// Training runs forward (activations stored as BF16 [sample][64]), the loss
// gradient per sample, backward through the layers (the transposed weights,
// same permutation trick), the hash-grid gradient by atomics, then the weight
// gradients as split-K GEMMs dW = dH^T H over the batch, staged through LDS.
#include "guide_math.h"
#include "accel_internal.h"

#include <math.h>
#include <string.h>

#include <new>
#include <random>
#include <vector>

using namespace ymx::guide;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

struct ymx_guide
{
	ymx_guide_params p;
	uint32_t table, n_grid;
	int32_t res[kLevels];
	float *grid, *grid_m, *grid_v, *grid_g;      // FP32 master, Adam moments, gradient
	uint16_t *grid_bf;                           // BF16 copy read by the kernels
	float *mlp, *mlp_m, *mlp_v, *mlp_g;          // W1 W2 W3 (64 x 64), W4 (48 x 64)
	uint16_t *w_bf;                              // 4 x [64 out][64 in] (W4 rows 48..63 zero)
	uint16_t *wt_bf;                             // 4 x [64 in][64 out] (W4^T columns 48..63 zero)
	uint16_t *X, *H1, *H2, *H3, *dO, *dH3, *dH2, *dH1;  // [max_batch][64]
	float *dXg;                                  // [level][max_batch] float2 grid-feature gradients (dense levels)
	uint32_t slab_mask;                          // dense levels whose gradient is summed per slab (k_grid_dense)
	uint32_t dense_off[kLevels + 1];             // prefix of the slab levels' entry counts
	uint32_t used_off[kLevels + 1];              // prefix of the used entries (dense: (res + 1)^3, hashed: table)
	float *dgrad_part;                           // [kGridParts][dense entries] float2 partial sums
	float *wgrad_part;                           // [kWgradBlocks][4][64 x 64] partial weight gradients
	double *loss;
	int step;
	float *selection;                            // learned selection probability of the path tracer (NULL until asked)
};

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMlp = YMX_GUIDE_MLP_WEIGHTS;
constexpr int kWgradSamples = 64;  // samples per LDS stage of the weight-gradient GEMM
constexpr uint32_t kSlabEntries = 8192;  // lattice entries per block of the dense-level gradient (64 KiB of LDS)
constexpr int kMaxSlabs = 96;
constexpr int kGridParts = 16;             // slices of the batch per slab, summed by k_grid_reduce
constexpr int kGridBlock = 1024;           // lanes of a k_grid_dense block (2 blocks of 64 KiB LDS per CU)
constexpr int kWgradSlices = 8;            // k_wgrad_reduce: slices of the partials, summed by atomics
constexpr int kWgradBlocks = 128;          // slices of the batch of the weight-gradient GEMMs

struct GArgs
{
	const uint16_t *grid_bf;
	const uint16_t *w_bf, *wt_bf;
	float *grid_g, *mlp_g;
	uint16_t *X, *H1, *H2, *H3, *dO, *dH3, *dH2, *dH1;
	float *dXg;
	uint32_t slab_mask, dxg_stride;
	double *loss;
	uint32_t table_mask;
	int32_t res[kLevels];
	float bmin[3], inv_ext[3];
};

union Frag
{
	bf16x8 v;
	uint16_t h[8];
	uint4 u;
	uint2 d[2];
};

__device__ __forceinline__ f32x4 mfma(const Frag &a, const Frag &b, f32x4 c)
{
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t count_of(uint32_t n, const uint32_t *n_dev)
{
	if(!n_dev) return n;
	const uint32_t m = *n_dev;
	return m < n ? m : n;
}

// Position of a sample on the unit cube of the scene box.
__device__ __forceinline__ void unit_pos(const GArgs &g, const ymx_guide_sample &r, float p[3])
{
	for(int a = 0; a < 3; ++a)
	{
		float v = (r.pos[a] - g.bmin[a]) * g.inv_ext[a];
		p[a] = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
	}
}

// Trilinear corners of one level: entry index and weight of corner c (bit 0 x, 1 y, 2 z).
struct Corners
{
	uint32_t idx[8];
	float w[8];
};
__device__ __forceinline__ void level_corners(const GArgs &g, int lev, const float p[3], Corners &c)
{
	const int r = g.res[lev];
	uint32_t b[3];
	float f[3];
	for(int a = 0; a < 3; ++a)
	{
		const float s = p[a] * (float)r;
		const float fl = floorf(s);
		b[a] = (uint32_t)fl;
		f[a] = s - fl;
	}
	for(int k = 0; k < 8; ++k)
	{
		const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
		const float wx = dx ? f[0] : 1.f - f[0];
		const float wy = dy ? f[1] : 1.f - f[1];
		const float wz = dz ? f[2] : 1.f - f[2];
		c.w[k] = (wx * wy) * wz;
		c.idx[k] = grid_index(r, g.table_mask, b[0] + dx, b[1] + dy, b[2] + dz);
	}
}

// Two consecutive lattice entries (2 BF16 features each) in one 8-byte load;
// entries are only 4-byte aligned, which global loads accept.
typedef unsigned long long __attribute__((aligned(4))) entry_pair_t;

// Interpolated features of one level.  On a dense level the corners (x, y, z)
// and (x + 1, y, z) are adjacent entries, so the 8 corners take 4 paired loads;
// at x == res (only on the upper face, where the x weight of the upper corner
// is 0) the pair starts at res - 1 with weights (0, 1), which sums the same
// terms.  The products and the order of the sum are those of level_corners.
__device__ __forceinline__ void level_features(const GArgs &g, int lev, const float p[3], float &f0, float &f1)
{
	const int r = g.res[lev];
	const uint32_t table = g.table_mask + 1u;
	const uint32_t *lv = reinterpret_cast<const uint32_t *>(g.grid_bf) + (size_t)lev * table;
	const uint32_t r1 = (uint32_t)r + 1u;
	f0 = 0.f;
	f1 = 0.f;
	if((uint64_t)r1 * r1 * r1 > (uint64_t)table)
	{
		Corners c;
		level_corners(g, lev, p, c);
		for(int k = 0; k < 8; ++k)
		{
			const uint32_t pair = lv[c.idx[k]];
			f0 = f0 + c.w[k] * bf2f((uint16_t)(pair & 0xFFFFu));
			f1 = f1 + c.w[k] * bf2f((uint16_t)(pair >> 16));
		}
		return;
	}
	uint32_t b[3];
	float f[3];
	for(int a = 0; a < 3; ++a)
	{
		const float sc = p[a] * (float)r;
		const float fl = floorf(sc);
		b[a] = (uint32_t)fl;
		f[a] = sc - fl;
	}
	const uint32_t ur = (uint32_t)r;
	if(b[0] >= ur)
	{
		b[0] = ur - 1u;  // res >= 2
		f[0] = 1.f;
	}
	for(int k = 0; k < 8; k += 2)
	{
		const uint32_t dy = (k >> 1) & 1, dz = (k >> 2) & 1;
		const uint32_t y = b[1] + dy < ur ? b[1] + dy : ur;
		const uint32_t z = b[2] + dz < ur ? b[2] + dz : ur;
		const float wy = dy ? f[1] : 1.f - f[1];
		const float wz = dz ? f[2] : 1.f - f[2];
		const float w0 = ((1.f - f[0]) * wy) * wz;
		const float w1 = (f[0] * wy) * wz;
		const unsigned long long e2 = *reinterpret_cast<const entry_pair_t *>(lv + b[0] + y * r1 + z * r1 * r1);
		const uint32_t e0 = (uint32_t)e2, e1 = (uint32_t)(e2 >> 32);
		f0 = f0 + w0 * bf2f((uint16_t)(e0 & 0xFFFFu));
		f1 = f1 + w0 * bf2f((uint16_t)(e0 >> 16));
		f0 = f0 + w1 * bf2f((uint16_t)(e1 & 0xFFFFu));
		f1 = f1 + w1 * bf2f((uint16_t)(e1 >> 16));
	}
}

// First-layer input of the lane: K-block 0 holds levels 4g .. 4g+3 (2 features
// each), K-block 1 the normal, wo and a constant 1 (lane group 0 only).
__device__ __forceinline__ void encode(const GArgs &g, const ymx_guide_sample &r, int grp, Frag &x0, Frag &x1)
{
	float p[3];
	unit_pos(g, r, p);
	for(int q = 0; q < 4; ++q)
	{
		float f0, f1;
		level_features(g, 4 * grp + q, p, f0, f1);
		x0.h[2 * q] = f2bf(f0);
		x0.h[2 * q + 1] = f2bf(f1);
	}
	for(int j = 0; j < 8; ++j) x1.h[j] = 0;
	if(grp == 0)
	{
		x1.h[0] = f2bf(r.normal[0]);
		x1.h[1] = f2bf(r.normal[1]);
		x1.h[2] = f2bf(r.normal[2]);
		x1.h[3] = f2bf(r.wo[0]);
		x1.h[4] = f2bf(r.wo[1]);
		x1.h[5] = f2bf(r.wo[2]);
		x1.h[6] = f2bf(1.f);
	}
}

// A operand of a layer whose input arrives in the permuted order (see the
// file comment): row `row` of the [64][64] matrix m, K-step ks, lane group grp.
__device__ __forceinline__ Frag load_a_perm(const uint16_t *m, int row, int ks, int grp)
{
	Frag a;
	const uint16_t *r = m + row * 64 + 32 * ks;
	a.d[0] = *reinterpret_cast<const uint2 *>(r + 4 * grp);
	a.d[1] = *reinterpret_cast<const uint2 *>(r + 16 + 4 * grp);
	return a;
}
__device__ __forceinline__ Frag load_a(const uint16_t *m, int row, int ks, int grp)
{
	Frag a;
	a.u = *reinterpret_cast<const uint4 *>(m + row * 64 + 32 * ks + 8 * grp);
	return a;
}

// ReLU'd accumulators of tiles 2ks, 2ks+1 as the next B operand (BF16).
__device__ __forceinline__ Frag acc_to_b(const f32x4 *acc, int ks, bool relu)
{
	Frag b;
	for(int i = 0; i < 4; ++i)
	{
		float v0 = acc[2 * ks][i], v1 = acc[2 * ks + 1][i];
		if(relu)
		{
			v0 = v0 > 0.f ? v0 : 0.f;
			v1 = v1 > 0.f ? v1 : 0.f;
		}
		b.h[i] = f2bf(v0);
		b.h[4 + i] = f2bf(v1);
	}
	return b;
}

// Store the lane's part of a [sample][64] activation row (positions of acc_to_b).
__device__ __forceinline__ void store_act(uint16_t *buf, uint32_t s, int grp, const Frag *b)
{
	uint16_t *row = buf + (size_t)s * 64;
	for(int ks = 0; ks < 2; ++ks)
	{
		*reinterpret_cast<uint2 *>(row + 32 * ks + 4 * grp) = b[ks].d[0];
		*reinterpret_cast<uint2 *>(row + 32 * ks + 16 + 4 * grp) = b[ks].d[1];
	}
}

// One hidden layer: 4 output tiles, 2 K-steps.
__device__ __forceinline__ void layer(const uint16_t *w, const Frag *in, bool perm, int sl, int grp, f32x4 *acc,
                                      int tiles)
{
	for(int t = 0; t < tiles; ++t)
	{
		acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
		for(int ks = 0; ks < 2; ++ks)
		{
			const Frag a = perm ? load_a_perm(w, 16 * t + sl, ks, grp) : load_a(w, 16 * t + sl, ks, grp);
			acc[t] = mfma(a, in[ks], acc[t]);
		}
	}
}

// Mixture decoding of 40 raw outputs (unit mu, clamped log kappa, softmax weights).
__device__ void decode(const float *raw, float *mix)
{
	float mx = -INFINITY;
	for(int k = 0; k < kLobes; ++k) mx = fmaxf(mx, raw[5 * k + 4]);
	float e[kLobes], sum = 0.f;
	for(int k = 0; k < kLobes; ++k)
	{
		e[k] = expf(raw[5 * k + 4] - mx);
		sum += e[k];
	}
	for(int k = 0; k < kLobes; ++k)
	{
		const float *l = raw + 5 * k;
		float n = sqrtf(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
		const float inv = 1.f / fmaxf(n, 1e-8f);
		const float s = fminf(fmaxf(l[3], kKappaLogMin), kKappaLogMax);
		const float kappa = expf(s);
		float *o = mix + 6 * k;
		o[0] = l[0] * inv;
		o[1] = l[1] * inv;
		o[2] = l[2] * inv;
		o[3] = kappa;
		o[4] = e[k] / sum;
		o[5] = kappa / (6.283185307179586f * (-expm1f(-2.f * kappa)));
	}
}

// decode() of lobes 2 grp and 2 grp + 1 only (12 floats to out), with the
// softmax max and sum taken over all 8 lobes in decode()'s order.
__device__ __forceinline__ void decode_pair(const float *raw, int grp, float *out)
{
	float mx = -INFINITY;
	for(int k = 0; k < kLobes; ++k) mx = fmaxf(mx, raw[5 * k + 4]);
	float sum = 0.f;
	for(int k = 0; k < kLobes; ++k) sum += expf(raw[5 * k + 4] - mx);
	float o[12];
	for(int j = 0; j < 2; ++j)
	{
		const int k = 2 * grp + j;
		const float *l = raw + 5 * k;
		const float n = sqrtf(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
		const float inv = 1.f / fmaxf(n, 1e-8f);
		const float sc = fminf(fmaxf(l[3], kKappaLogMin), kKappaLogMax);
		const float kappa = expf(sc);
		float *q = o + 6 * j;
		q[0] = l[0] * inv;
		q[1] = l[1] * inv;
		q[2] = l[2] * inv;
		q[3] = kappa;
		q[4] = expf(l[4] - mx) / sum;
		q[5] = kappa / (6.283185307179586f * (-expm1f(-2.f * kappa)));
	}
	float4 *dst = reinterpret_cast<float4 *>(out);
	for(int k = 0; k < 3; ++k) dst[k] = make_float4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
}

// d(-c log q)/d raw for one training sample (oracle/ora_guide.py dloss_dout).
__device__ float loss_grad(const float *raw, const ymx_guide_sample &r, float *d)
{
	float mix[kMix];
	decode(raw, mix);
	const float wx = r.w[0], wy = r.w[1], wz = r.w[2];
	float v[kLobes], cosv[kLobes], q = 0.f;
	for(int k = 0; k < kLobes; ++k)
	{
		const float *l = mix + 6 * k;
		cosv[k] = l[0] * wx + l[1] * wy + l[2] * wz;
		v[k] = l[4] * l[5] * expf(l[3] * (cosv[k] - 1.f));
		q += v[k];
	}
	const float c = r.c;
	const float qs = fmaxf(q, 1e-30f);
	for(int k = 0; k < kLobes; ++k)
	{
		const float *l = mix + 6 * k;
		const float *rk = raw + 5 * k;
		const float resp = v[k] / qs;
		const float kappa = l[3];
		d[5 * k + 4] = -c * (resp - l[4]);
		const float e2 = expf(-2.f * kappa);
		const float dk = 1.f / kappa - 2.f * e2 / (-expm1f(-2.f * kappa)) + cosv[k] - 1.f;
		const bool inside = rk[3] > kKappaLogMin && rk[3] < kKappaLogMax;
		d[5 * k + 3] = inside ? -c * resp * kappa * dk : 0.f;
		const float n = sqrtf(rk[0] * rk[0] + rk[1] * rk[1] + rk[2] * rk[2]);
		const float sc = -c * resp * kappa / fmaxf(n, 1e-8f);
		d[5 * k + 0] = sc * (wx - cosv[k] * l[0]);
		d[5 * k + 1] = sc * (wy - cosv[k] * l[1]);
		d[5 * k + 2] = sc * (wz - cosv[k] * l[2]);
	}
	for(int k = 5 * kLobes; k < 64; ++k) d[k] = 0.f;
	return -c * logf(qs);
}

// Training forward: the BF16 activations of every layer and the loss gradient.
__global__ __launch_bounds__(kBlock) void k_forward(GArgs g, const ymx_guide_sample *smp, uint32_t cap,
                                                     const uint32_t *n_dev)
{
	__shared__ float outs[kWaves][16][kOut];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int sl = lane & 15, grp = lane >> 4;
	const uint32_t n = count_of(cap, n_dev);
	const uint32_t tile0 = (blockIdx.x * kWaves + wave) * 16;
	const bool wave_active = tile0 < n;  // uniform over the wave: MFMAs run with all lanes
	const uint32_t s = tile0 + sl;
	const bool valid = s < n;
	if(wave_active)
	{
		ymx_guide_sample r;
		if(valid) r = smp[s];
		else memset(&r, 0, sizeof r);
		Frag in[2];
		encode(g, r, grp, in[0], in[1]);
		if(valid)
		{
			uint16_t *row = g.X + (size_t)s * 64;
			*reinterpret_cast<uint4 *>(row + 8 * grp) = in[0].u;
			*reinterpret_cast<uint4 *>(row + 32 + 8 * grp) = in[1].u;
		}
		f32x4 acc[4];
		layer(g.w_bf, in, false, sl, grp, acc, 4);
		uint16_t *hbuf[3] = {g.H1, g.H2, g.H3};
		for(int l = 1; l < 4; ++l)
		{
			in[0] = acc_to_b(acc, 0, true);
			in[1] = acc_to_b(acc, 1, true);
			if(valid) store_act(hbuf[l - 1], s, grp, in);
			layer(g.w_bf + l * 4096, in, true, sl, grp, acc, l < 3 ? 4 : 3);
		}
		for(int t = 0; t < 3; ++t)
			for(int i = 0; i < 4; ++i) outs[wave][sl][16 * t + 4 * grp + i] = acc[t][i];
	}
	__syncthreads();
	if(!wave_active || grp != 0 || !valid) return;
	const ymx_guide_sample r = smp[s];
	const float *raw = outs[wave][sl];
	float d[64];
	const float l = loss_grad(raw, r, d);
	atomicAdd(g.loss, (double)l);
	uint16_t *row = g.dO + (size_t)s * 64;
	for(int k = 0; k < 64; k += 8)
	{
		Frag f;
		for(int j = 0; j < 8; ++j) f.h[j] = f2bf(d[k + j]);
		*reinterpret_cast<uint4 *>(row + k) = f.u;
	}
}

// Inference, persistent: each block stages the four weight matrices in LDS
// once (32 KiB) and its waves then take 16-query tiles until the batch is
// done.  Every A operand is one 16-byte LDS read: row r of layer l is 8 chunks
// of 8 BF16, chunk c = 4 ks + grp holding the elements lane group grp needs
// at K-step ks (natural order for layer 1, the permuted order of the file
// comment for the others), stored at position c ^ (r & 7) so that the 16 rows
// a wave reads together fall on distinct banks.
constexpr int kInferWaves = 12;
constexpr int kInferBlock = 64 * kInferWaves;
constexpr int kInferBlocksPerCu = 2;

__device__ __forceinline__ void layer_lds(const uint4 *wl, int l, const Frag *in, int sl, int grp, f32x4 *acc, int tiles)
{
	for(int t = 0; t < tiles; ++t)
	{
		acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
		const int row = 16 * t + sl;
		for(int ks = 0; ks < 2; ++ks)
		{
			Frag a;
			a.u = wl[(l * 64 + row) * 8 + ((4 * ks + grp) ^ (row & 7))];
			acc[t] = mfma(a, in[ks], acc[t]);
		}
	}
}

__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kInferBlock) void k_infer(GArgs g, const ymx_guide_sample *smp, uint32_t cap,
                                                        const uint32_t *n_dev, float *mix_out)
{
	__shared__ uint4 wl[4 * 64 * 8];
	__shared__ float outs[kInferWaves][16][kOut];
	for(int e = threadIdx.x; e < 4 * 64 * 8; e += kInferBlock)
	{
		const int l = e >> 9, row = (e >> 3) & 63, c = e & 7;
		const uint16_t *src = g.w_bf + l * 4096 + row * 64;
		Frag f;
		if(l == 0)
			f.u = *reinterpret_cast<const uint4 *>(src + 8 * c);
		else
		{
			const int ks = c >> 2, grp = c & 3;
			f.d[0] = *reinterpret_cast<const uint2 *>(src + 32 * ks + 4 * grp);
			f.d[1] = *reinterpret_cast<const uint2 *>(src + 32 * ks + 16 + 4 * grp);
		}
		wl[(l * 64 + row) * 8 + (c ^ (row & 7))] = f.u;
	}
	__syncthreads();
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int sl = lane & 15, grp = lane >> 4;
	const uint32_t n = count_of(cap, n_dev);
	const uint32_t tiles = (n + 15u) / 16u;
	for(uint32_t tile = blockIdx.x * kInferWaves + wave; tile < tiles; tile += gridDim.x * kInferWaves)
	{
		const uint32_t s = tile * 16u + sl;
		const bool valid = s < n;
		ymx_guide_sample r;
		if(valid) r = smp[s];
		else memset(&r, 0, sizeof r);
		Frag in[2];
		encode(g, r, grp, in[0], in[1]);
		f32x4 acc[4];
		layer_lds(wl, 0, in, sl, grp, acc, 4);
		for(int l = 1; l < 4; ++l)
		{
			in[0] = acc_to_b(acc, 0, true);
			in[1] = acc_to_b(acc, 1, true);
			layer_lds(wl, l, in, sl, grp, acc, l < 3 ? 4 : 3);
		}
		for(int t = 0; t < 3; ++t)
			for(int i = 0; i < 4; ++i) outs[wave][sl][16 * t + 4 * grp + i] = acc[t][i];
		wave_sync();
		if(valid) decode_pair(outs[wave][sl], grp, mix_out + (size_t)s * kMix + 12 * grp);
		wave_sync();
	}
}

// Backward through the layers: dH (BF16, ReLU mask applied) of layers 3, 2, 1
// and the hash-grid gradient.
__global__ __launch_bounds__(kBlock) void k_backward(GArgs g, const ymx_guide_sample *smp, uint32_t cap,
                                                      const uint32_t *n_dev)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int sl = lane & 15, grp = lane >> 4;
	const uint32_t n = count_of(cap, n_dev);
	const uint32_t tile0 = (blockIdx.x * kWaves + wave) * 16;
	if(tile0 >= n) return;
	const uint32_t s = tile0 + sl;
	const bool valid = s < n;
	const uint32_t sr = valid ? s : tile0;  // rows of an idle lane are read but not used
	Frag in[2];
	in[0].u = *reinterpret_cast<const uint4 *>(g.dO + (size_t)sr * 64 + 8 * grp);
	in[1].u = *reinterpret_cast<const uint4 *>(g.dO + (size_t)sr * 64 + 32 + 8 * grp);
	if(!valid)
	{
		in[0].u = make_uint4(0, 0, 0, 0);
		in[1].u = make_uint4(0, 0, 0, 0);
	}
	f32x4 acc[4];
	// dH3 = W4^T dO (input positions of dO in natural order)
	layer(g.wt_bf + 3 * 4096, in, false, sl, grp, acc, 4);
	const uint16_t *hact[3] = {g.H3, g.H2, g.H1};
	uint16_t *dbuf[3] = {g.dH3, g.dH2, g.dH1};
	for(int l = 0; l < 3; ++l)
	{
		// ReLU mask from the stored activation at the accumulator positions
		const uint16_t *hrow = hact[l] + (size_t)sr * 64;
		for(int t = 0; t < 4; ++t)
		{
			const uint2 hv = *reinterpret_cast<const uint2 *>(hrow + 16 * t + 4 * grp);
			const uint16_t hh[4] = {(uint16_t)(hv.x & 0xFFFFu), (uint16_t)(hv.x >> 16), (uint16_t)(hv.y & 0xFFFFu),
			                        (uint16_t)(hv.y >> 16)};
			for(int i = 0; i < 4; ++i)
				if(!(bf2f(hh[i]) > 0.f)) acc[t][i] = 0.f;
		}
		in[0] = acc_to_b(acc, 0, false);
		in[1] = acc_to_b(acc, 1, false);
		if(valid) store_act(dbuf[l], s, grp, in);
		// next: dH_{l-1} = W_l^T dH_l (layer 3 - l), input positions permuted
		layer(g.wt_bf + (2 - l) * 4096, in, true, sl, grp, acc, l < 2 ? 4 : 2);
	}
	if(!valid) return;
	// acc[t][i] = dX[16 t + 4 grp + i], t = 0, 1: the grid features of levels
	// 8t + 2 grp and 8t + 2 grp + 1
	const ymx_guide_sample r = smp[s];
	float p[3];
	unit_pos(g, r, p);
	const uint32_t table = g.table_mask + 1u;
	for(int t = 0; t < 2; ++t)
		for(int q = 0; q < 2; ++q)
		{
			const int lev = 8 * t + 2 * grp + q;
			const float d0 = acc[t][2 * q], d1 = acc[t][2 * q + 1];
			if((g.slab_mask >> lev) & 1u)
			{
				// dense level: summed per lattice slab in LDS by k_grid_dense
				*reinterpret_cast<float2 *>(g.dXg + 2 * ((size_t)lev * g.dxg_stride + s)) = make_float2(d0, d1);
				continue;
			}
			Corners c;
			level_corners(g, lev, p, c);
			float *lv = g.grid_g + (size_t)lev * table * kFeat;
			for(int k = 0; k < 8; ++k)
			{
				atomicAdd(lv + 2 * (size_t)c.idx[k], c.w[k] * d0);
				atomicAdd(lv + 2 * (size_t)c.idx[k] + 1, c.w[k] * d1);
			}
		}
}

// Grid gradient of the dense levels without global atomics.  A slab is a run
// of whole z-planes of one level (at most kSlabEntries entries); block (slab,
// part) adds the corner contributions of its slice of the batch that fall in
// the slab into LDS (corner weights and the sum order within a sample as in
// level_corners) and stores the slab to its part of dgrad_part; k_grid_reduce
// then sums the parts of every entry into the gradient.
struct Slab
{
	int32_t lev;
	uint32_t z0, z1, off;  // planes [z0, z1); off: first entry in the dense entry space
};
struct SlabList
{
	Slab s[kMaxSlabs];
	int32_t count;
};
__global__ __launch_bounds__(kGridBlock) void k_grid_dense(GArgs g, SlabList sl, const ymx_guide_sample *smp,
                                                        uint32_t cap, const uint32_t *n_dev, float *part_out,
                                                        uint32_t dense_total)
{
	extern __shared__ float lds_slab[];
	const Slab sb = sl.s[blockIdx.x];
	const uint32_t part = blockIdx.y;
	const uint32_t r = (uint32_t)g.res[sb.lev], r1 = r + 1u, plane = r1 * r1;
	const uint32_t entries = (sb.z1 - sb.z0) * plane;  // <= kSlabEntries, checked on the host
	for(uint32_t e = threadIdx.x; e < 2 * entries; e += kGridBlock) lds_slab[e] = 0.f;
	__syncthreads();
	const uint32_t n = count_of(cap, n_dev);
	const uint32_t per = (n + kGridParts - 1) / kGridParts;
	const uint32_t s0 = part * per, s1 = s0 + per < n ? s0 + per : n;
	const float2 *dxg = reinterpret_cast<const float2 *>(g.dXg) + (size_t)sb.lev * g.dxg_stride;
	for(uint32_t s = s0 + threadIdx.x; s < s1; s += kGridBlock)
	{
		float p[3];
		unit_pos(g, smp[s], p);
		uint32_t b[3];
		float f[3];
		for(int a = 0; a < 3; ++a)
		{
			const float sc = p[a] * (float)r;
			const float fl = floorf(sc);
			b[a] = (uint32_t)fl;
			f[a] = sc - fl;
		}
		if(b[2] + 1u < sb.z0 || b[2] >= sb.z1) continue;
		const float2 d = dxg[s];
		for(int k = 0; k < 8; ++k)
		{
			const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
			const uint32_t z = b[2] + dz < r ? b[2] + dz : r;
			if(z < sb.z0 || z >= sb.z1) continue;
			const uint32_t x = b[0] + dx < r ? b[0] + dx : r;
			const uint32_t y = b[1] + dy < r ? b[1] + dy : r;
			const float wx = dx ? f[0] : 1.f - f[0];
			const float wy = dy ? f[1] : 1.f - f[1];
			const float wz = dz ? f[2] : 1.f - f[2];
			const float w = (wx * wy) * wz;
			const uint32_t e = x + y * r1 + (z - sb.z0) * plane;
			atomicAdd(&lds_slab[2 * e], w * d.x);
			atomicAdd(&lds_slab[2 * e + 1], w * d.y);
		}
	}
	__syncthreads();
	float *dst = part_out + 2 * ((size_t)part * dense_total + sb.off);
	for(uint32_t e = threadIdx.x; e < 2 * entries; e += kGridBlock) dst[e] = lds_slab[e];
}

struct LevelOffsets
{
	uint32_t off[kLevels + 1];  // prefix of per-level entry counts in a compact entry space
	uint32_t mask;              // levels present in that space
};
__device__ __forceinline__ int level_of(const LevelOffsets &lo, uint32_t e)
{
	int l = 0;
	while(l < kLevels - 1 && e >= lo.off[l + 1]) ++l;
	return l;
}

__global__ void k_grid_reduce(const float *part, uint32_t dense_total, LevelOffsets lo, float *grid_g, uint32_t table)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= 2 * dense_total) return;
	float acc = 0.f;
	for(int k = 0; k < kGridParts; ++k) acc += part[(size_t)k * 2 * dense_total + i];
	const uint32_t e = i >> 1;
	const int l = level_of(lo, e);
	grid_g[2 * ((size_t)l * table + (e - lo.off[l])) + (i & 1u)] = acc;
}

// dW (64 x 64 FP32) += sum over samples of dH[s][o] * H[s][i], split over
// blocks along the samples.  Wave w computes output rows 16w .. 16w+15.
struct WgradLayers
{
	const uint16_t *dH[4], *H[4];
};
__global__ __launch_bounds__(kBlock) void k_wgrad(WgradLayers wl, uint32_t cap, const uint32_t *n_dev, float *part)
{
	const int layer = blockIdx.y;
	const uint16_t *dH = wl.dH[layer], *H = wl.H[layer];
	const int rows = layer < 3 ? 64 : 48;
	float *dW = part + ((size_t)blockIdx.x * 4 + layer) * 4096;
	// [feature][sample], rows padded to 144 B (a multiple of 16 for 128-bit reads)
	__shared__ __attribute__((aligned(16))) uint16_t sd[64][kWgradSamples + 8];
	__shared__ __attribute__((aligned(16))) uint16_t sh[64][kWgradSamples + 8];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int sl = lane & 15, grp = lane >> 4;
	const uint32_t n = count_of(cap, n_dev);
	const uint32_t per = ((n + gridDim.x - 1) / gridDim.x + kWgradSamples - 1) / kWgradSamples * kWgradSamples;
	const uint32_t b0 = blockIdx.x * per;
	const uint32_t b1 = b0 + per < n ? b0 + per : n;
	f32x4 acc[4] = {};
	for(uint32_t base = b0; base < b1; base += kWgradSamples)
	{
		// stage 64 samples x 64 features of both operands, transposed
		for(int e = threadIdx.x; e < kWgradSamples * 8; e += kBlock)
		{
			const int sm = e >> 3, part = e & 7;
			const uint32_t s = base + sm;
			Frag a, h;
			if(s < b1)
			{
				a.u = *reinterpret_cast<const uint4 *>(dH + (size_t)s * 64 + 8 * part);
				h.u = *reinterpret_cast<const uint4 *>(H + (size_t)s * 64 + 8 * part);
			}
			else
			{
				a.u = make_uint4(0, 0, 0, 0);
				h.u = make_uint4(0, 0, 0, 0);
			}
			for(int j = 0; j < 8; ++j)
			{
				sd[8 * part + j][sm] = a.h[j];
				sh[8 * part + j][sm] = h.h[j];
			}
		}
		__syncthreads();
		for(int kb = 0; kb < kWgradSamples / 32; ++kb)
		{
			Frag a, b;
			a.u = *reinterpret_cast<const uint4 *>(&sd[16 * wave + sl][32 * kb + 8 * grp]);
			for(int t = 0; t < 4; ++t)
			{
				b.u = *reinterpret_cast<const uint4 *>(&sh[16 * t + sl][32 * kb + 8 * grp]);
				acc[t] = mfma(a, b, acc[t]);
			}
		}
		__syncthreads();
	}
	// acc[t][i] = dW[16 wave + 4 grp + i][16 t + sl]
	for(int t = 0; t < 4; ++t)
		for(int i = 0; i < 4; ++i)
		{
			const int o = 16 * wave + 4 * grp + i;
			dW[o * 64 + 16 * t + sl] = o < rows ? acc[t][i] : 0.f;
		}
}

// mlp_g[i] += sum over a slice (blockIdx.y) of the k_wgrad blocks' partial
// gradients; mlp_g is cleared before.
__global__ void k_wgrad_reduce(const float *part, uint32_t blocks, float *mlp_g)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= 4 * 4096) return;
	const uint32_t per = (blocks + kWgradSlices - 1) / kWgradSlices;
	const uint32_t b0 = blockIdx.y * per, b1 = b0 + per < blocks ? b0 + per : blocks;
	float acc = 0.f;
	for(uint32_t b = b0; b < b1; ++b) acc += part[(size_t)b * 4 * 4096 + i];
	atomicAdd(mlp_g + i, acc);
}

// Adam on FP32 master weights (Kingma & Ba 2015, bias-corrected), then the
// BF16 copy.
__global__ void k_adam(float *p, float *m, float *v, float *gr, uint16_t *pb, uint32_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) return;
	const float gi = gr[i];
	const float mi = b1 * m[i] + (1.f - b1) * gi;
	const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
	p[i] = pi;
	gr[i] = 0.f;
	if(pb) pb[i] = f2bf(pi);
}

// Adam over the lattice entries in use (a dense level uses its first
// (res + 1)^3 entries; an entry no sample reaches keeps a zero gradient and
// zero moments, so Adam would leave it unchanged).
__global__ void k_adam_grid(float *p, float *m, float *v, float *gr, uint16_t *pb, LevelOffsets lo, uint32_t table,
                            float lr, float b1, float b2, float eps, float c1, float c2)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= 2 * lo.off[kLevels]) return;
	const uint32_t e = i >> 1;
	const int l = level_of(lo, e);
	const size_t k = 2 * ((size_t)l * table + (e - lo.off[l])) + (i & 1u);
	const float gi = gr[k];
	const float mi = b1 * m[k] + (1.f - b1) * gi;
	const float vi = b2 * v[k] + (1.f - b2) * gi * gi;
	m[k] = mi;
	v[k] = vi;
	const float pi = p[k] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
	p[k] = pi;
	gr[k] = 0.f;
	pb[k] = f2bf(pi);
}

// BF16 weight layouts from the FP32 master: W (rows padded to 64) and W^T.
__global__ void k_weights_bf(const float *mlp, uint16_t *w_bf, uint16_t *wt_bf)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= 4 * 4096) return;
	const int l = i / 4096, o = (i % 4096) / 64, c = i % 64;
	const int rows = l < 3 ? 64 : 48;
	const float v = o < rows ? mlp[l * 4096 + o * 64 + c] : 0.f;
	w_bf[l * 4096 + o * 64 + c] = f2bf(v);
	wt_bf[l * 4096 + c * 64 + o] = f2bf(v);
}

GArgs args_of(const ymx_guide *g)
{
	GArgs a;
	a.grid_bf = g->grid_bf;
	a.w_bf = g->w_bf;
	a.wt_bf = g->wt_bf;
	a.grid_g = g->grid_g;
	a.mlp_g = g->mlp_g;
	a.X = g->X;
	a.H1 = g->H1;
	a.H2 = g->H2;
	a.H3 = g->H3;
	a.dO = g->dO;
	a.dH3 = g->dH3;
	a.dH2 = g->dH2;
	a.dH1 = g->dH1;
	a.dXg = g->dXg;
	a.slab_mask = g->slab_mask;
	a.dxg_stride = g->p.max_batch;
	a.loss = g->loss;
	a.table_mask = g->table - 1u;
	for(int l = 0; l < kLevels; ++l) a.res[l] = g->res[l];
	for(int k = 0; k < 3; ++k)
	{
		a.bmin[k] = g->p.bmin[k];
		a.inv_ext[k] = 1.f / (g->p.bmax[k] - g->p.bmin[k]);
	}
	return a;
}

void free_guide(ymx_guide *g)
{
	if(!g) return;
	void *ptrs[] = {g->grid, g->grid_m, g->grid_v, g->grid_g, g->grid_bf, g->mlp, g->mlp_m, g->mlp_v, g->mlp_g,
	                g->w_bf, g->wt_bf, g->X, g->H1, g->H2, g->H3, g->dO, g->dH3, g->dH2, g->dH1, g->dXg, g->dgrad_part, g->wgrad_part, g->loss,
	                g->selection};
	for(void *p : ptrs)
		if(p) (void)hipFree(p);
	delete g;
}

template <typename T> hipError_t galloc(T *&p, size_t count) { return hipMalloc((void **)&p, sizeof(T) * (count ? count : 1)); }

int refresh_bf16(ymx_guide *g, hipStream_t s)
{
	k_weights_bf<<<(4 * 4096 + 255) / 256, 256, 0, s>>>(g->mlp, g->w_bf, g->wt_bf);
	YMX_HIP_CHECK(hipGetLastError());
	return YMX_OK;
}

}  // namespace

uint32_t ymx_guide_batch_limit(const ymx_guide *g) { return g ? g->p.max_batch : 0u; }

float *ymx_guide_selection_table(ymx_guide *g, size_t floats)
{
	if(!g) return nullptr;
	if(!g->selection)
	{
		// a new table: every logit and moment zero (probability 0.6 until learned)
		if(hipMalloc((void **)&g->selection, sizeof(float) * floats) != hipSuccess ||
		   hipMemset(g->selection, 0, sizeof(float) * floats) != hipSuccess)
		{
			if(g->selection) (void)hipFree(g->selection);
			g->selection = nullptr;
			return nullptr;
		}
	}
	return g->selection;
}

void ymx_guide_bounds(const ymx_guide *g, float bmin[3], float bmax[3])
{
	for(int k = 0; k < 3; ++k)
	{
		bmin[k] = g->p.bmin[k];
		bmax[k] = g->p.bmax[k];
	}
}

extern "C" {

int ymx_guide_create(const ymx_guide_params *p, ymx_guide **out)
{
	if(!p || !out || p->struct_size != sizeof(ymx_guide_params))
	{
		ymx_set_error("ymx_guide_create: NULL argument or struct_size mismatch");
		return YMX_ERR_INVALID_ARG;
	}
	*out = nullptr;
	if(p->log2_table_size < 8 || p->log2_table_size > 22 || p->n_min < 2 || p->n_max < p->n_min ||
	   p->n_max > (1 << 20) || p->max_batch == 0 || p->max_batch > (1u << 24) ||
	   !(p->bmax[0] > p->bmin[0] && p->bmax[1] > p->bmin[1] && p->bmax[2] > p->bmin[2]))
	{
		ymx_set_error("ymx_guide_create: unsupported parameters");
		return YMX_ERR_UNSUPPORTED;
	}
	ymx_guide *g = new(std::nothrow) ymx_guide();
	if(!g)
	{
		ymx_set_error("ymx_guide_create: host allocation failed");
		return YMX_ERR_OUT_OF_MEMORY;
	}
	memset(g, 0, sizeof *g);
	g->p = *p;
	g->table = 1u << p->log2_table_size;
	g->n_grid = kLevels * g->table * kFeat;
	// N_l = floor(N_min b^l), b = exp((ln N_max - ln N_min) / (L - 1))
	const double b = exp((log((double)p->n_max) - log((double)p->n_min)) / (kLevels - 1));
	for(int l = 0; l < kLevels; ++l) g->res[l] = (int32_t)floor((double)p->n_min * pow(b, (double)l));
#define G_TRY(expr)                                                            \
	do                                                                         \
	{                                                                          \
		hipError_t e_ = (expr);                                                \
		if(e_ != hipSuccess)                                                   \
		{                                                                      \
			ymx_set_error("%s: %s", #expr, hipGetErrorString(e_));             \
			free_guide(g);                                                     \
			return e_ == hipErrorOutOfMemory ? YMX_ERR_OUT_OF_MEMORY : YMX_ERR_HIP; \
		}                                                                      \
	} while(0)
	G_TRY(galloc(g->grid, g->n_grid));
	G_TRY(galloc(g->grid_m, g->n_grid));
	G_TRY(galloc(g->grid_v, g->n_grid));
	G_TRY(galloc(g->grid_g, g->n_grid));
	G_TRY(galloc(g->grid_bf, g->n_grid));
	G_TRY(galloc(g->mlp, kMlp));
	G_TRY(galloc(g->mlp_m, kMlp));
	G_TRY(galloc(g->mlp_v, kMlp));
	G_TRY(galloc(g->mlp_g, kMlp));
	G_TRY(galloc(g->w_bf, 4 * 4096));
	G_TRY(galloc(g->wt_bf, 4 * 4096));
	uint16_t **acts[] = {&g->X, &g->H1, &g->H2, &g->H3, &g->dO, &g->dH3, &g->dH2, &g->dH1};
	for(uint16_t **a : acts) G_TRY(galloc(*a, (size_t)p->max_batch * 64));
	// Dense levels whose z-planes fit a slab, coarse first while the slab list
	// has room, sum their gradient in LDS; the others use global atomics.
	g->slab_mask = 0;
	g->dense_off[0] = g->used_off[0] = 0;
	int slabs = 0;
	for(int l = 0; l < kLevels; ++l)
	{
		const uint64_t r1 = (uint64_t)g->res[l] + 1u;
		const bool dense = r1 * r1 * r1 <= (uint64_t)g->table;
		bool slab = false;
		if(dense && r1 * r1 <= kSlabEntries)
		{
			const uint64_t per = kSlabEntries / (r1 * r1);
			const int need = (int)((r1 + per - 1) / per);
			if(slabs + need <= kMaxSlabs)
			{
				slab = true;
				slabs += need;
				g->slab_mask |= 1u << l;
			}
		}
		g->dense_off[l + 1] = g->dense_off[l] + (slab ? (uint32_t)(r1 * r1 * r1) : 0u);
		g->used_off[l + 1] = g->used_off[l] + (dense ? (uint32_t)(r1 * r1 * r1) : g->table);
	}
	if(g->slab_mask)
	{
		G_TRY(galloc(g->dXg, (size_t)p->max_batch * 2 * kLevels));
		G_TRY(galloc(g->dgrad_part, (size_t)kGridParts * 2 * g->dense_off[kLevels]));
	}
	G_TRY(galloc(g->wgrad_part, (size_t)kWgradBlocks * 4 * 4096));
	G_TRY(galloc(g->loss, 1));
	// initial weights: grid U(-1e-4, 1e-4), W U(-s, s) with s = sqrt(6 / fan_in)
	std::mt19937 rng(p->seed);
	std::uniform_real_distribution<float> u(-1.f, 1.f);
	std::vector<float> grid(g->n_grid), mlp(kMlp, 0.f);
	for(float &x : grid) x = 1e-4f * u(rng);
	for(int l = 0; l < 4; ++l)
	{
		const int rows = l < 3 ? 64 : 48;
		const float s = sqrtf(6.f / 64.f);
		for(int i = 0; i < rows * 64; ++i) mlp[l * 4096 + i] = s * u(rng);
	}
	for(int i = 5 * kLobes * 64; i < 48 * 64; ++i) mlp[3 * 4096 + i] = 0.f;
	*out = g;
	const int rc = ymx_guide_set_weights(g, grid.data(), mlp.data());
	if(rc != YMX_OK)
	{
		free_guide(g);
		*out = nullptr;
	}
	return rc;
#undef G_TRY
}

int ymx_guide_destroy(ymx_guide *g)
{
	if(!g) return YMX_OK;
	(void)hipDeviceSynchronize();
	free_guide(g);
	return YMX_OK;
}

int ymx_guide_set_weights(ymx_guide *g, const float *grid, const float *mlp)
{
	if(!g || !grid || !mlp)
	{
		ymx_set_error("ymx_guide_set_weights: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipMemcpy(g->grid, grid, sizeof(float) * g->n_grid, hipMemcpyHostToDevice));
	YMX_HIP_CHECK(hipMemcpy(g->mlp, mlp, sizeof(float) * kMlp, hipMemcpyHostToDevice));
	std::vector<uint16_t> gb(g->n_grid);
	for(uint32_t i = 0; i < g->n_grid; ++i) gb[i] = f2bf(grid[i]);
	YMX_HIP_CHECK(hipMemcpy(g->grid_bf, gb.data(), sizeof(uint16_t) * g->n_grid, hipMemcpyHostToDevice));
	YMX_HIP_CHECK(hipMemset(g->grid_m, 0, sizeof(float) * g->n_grid));
	YMX_HIP_CHECK(hipMemset(g->grid_v, 0, sizeof(float) * g->n_grid));
	YMX_HIP_CHECK(hipMemset(g->grid_g, 0, sizeof(float) * g->n_grid));
	YMX_HIP_CHECK(hipMemset(g->mlp_m, 0, sizeof(float) * kMlp));
	YMX_HIP_CHECK(hipMemset(g->mlp_v, 0, sizeof(float) * kMlp));
	YMX_HIP_CHECK(hipMemset(g->mlp_g, 0, sizeof(float) * kMlp));
	g->step = 0;
	const int rc = refresh_bf16(g, nullptr);
	if(rc != YMX_OK) return rc;
	YMX_HIP_CHECK(hipDeviceSynchronize());
	return YMX_OK;
}

int ymx_guide_get_weights(ymx_guide *g, float *grid, float *mlp)
{
	if(!g || !grid || !mlp)
	{
		ymx_set_error("ymx_guide_get_weights: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	YMX_HIP_CHECK(hipMemcpy(grid, g->grid, sizeof(float) * g->n_grid, hipMemcpyDeviceToHost));
	YMX_HIP_CHECK(hipMemcpy(mlp, g->mlp, sizeof(float) * kMlp, hipMemcpyDeviceToHost));
	return YMX_OK;
}

int ymx_guide_infer(ymx_guide *g, const ymx_guide_sample *queries, float *mix, uint32_t n, const uint32_t *n_dev,
                    void *stream)
{
	if(!g || (n && (!queries || !mix)))
	{
		ymx_set_error("ymx_guide_infer: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	if(n == 0) return YMX_OK;
	static int cus = 0;
	if(cus == 0)
	{
		int dev = 0, c = 0;
		if(hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
			c = 256;
		cus = c;
	}
	const uint32_t tiles = (n + 15u) / 16u;
	const uint32_t want = (tiles + kInferWaves - 1) / kInferWaves;
	const uint32_t blocks = want < (uint32_t)(kInferBlocksPerCu * cus) ? want : (uint32_t)(kInferBlocksPerCu * cus);
	k_infer<<<blocks, kInferBlock, 0, (hipStream_t)stream>>>(args_of(g), queries, n, n_dev, mix);
	YMX_HIP_CHECK(hipGetLastError());
	return YMX_OK;
}

int ymx_guide_backward(ymx_guide *g, const ymx_guide_sample *samples, uint32_t n, const uint32_t *n_dev,
                       void *stream)
{
	if(!g || (n && !samples))
	{
		ymx_set_error("ymx_guide_backward: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	if(n > g->p.max_batch)
	{
		ymx_set_error("ymx_guide_backward: %u samples exceed max_batch %u", n, g->p.max_batch);
		return YMX_ERR_INVALID_ARG;
	}
	hipStream_t s = (hipStream_t)stream;
	// dense levels: k_grid_reduce stores their gradient; hashed ones are summed by atomics
	for(int l = 0; l < kLevels; ++l)
		if(!((g->slab_mask >> l) & 1u))
			YMX_HIP_CHECK(hipMemsetAsync(g->grid_g + (size_t)l * g->table * kFeat, 0, sizeof(float) * g->table * kFeat, s));
	YMX_HIP_CHECK(hipMemsetAsync(g->loss, 0, sizeof(double), s));
	const GArgs a = args_of(g);
	if(n == 0)
	{
		for(int l = 0; l < kLevels; ++l)
			if((g->slab_mask >> l) & 1u)
				YMX_HIP_CHECK(hipMemsetAsync(g->grid_g + (size_t)l * g->table * kFeat, 0,
				                             sizeof(float) * (g->dense_off[l + 1] - g->dense_off[l]) * kFeat, s));
		YMX_HIP_CHECK(hipMemsetAsync(g->mlp_g, 0, sizeof(float) * kMlp, s));
		return YMX_OK;
	}
	const unsigned blocks = (n + 16 * kWaves - 1) / (16 * kWaves);
	k_forward<<<blocks, kBlock, 0, s>>>(a, samples, n, n_dev);
	YMX_HIP_CHECK(hipGetLastError());
	k_backward<<<blocks, kBlock, 0, s>>>(a, samples, n, n_dev);
	YMX_HIP_CHECK(hipGetLastError());
	if(g->slab_mask)
	{
		SlabList sl;
		sl.count = 0;
		for(int l = 0; l < kLevels; ++l)
		{
			if(!((g->slab_mask >> l) & 1u)) continue;
			const uint32_t r1 = (uint32_t)g->res[l] + 1u, plane = r1 * r1;
			const uint32_t per = kSlabEntries / plane;  // planes per slab (>= 1 for a slab level)
			for(uint32_t z0 = 0; z0 < r1; z0 += per)
			{
				if(sl.count == kMaxSlabs)
				{
					ymx_set_error("ymx_guide_backward: more than %d lattice slabs", kMaxSlabs);
					return YMX_ERR_UNSUPPORTED;
				}
				Slab &sb = sl.s[sl.count++];
				sb.lev = l;
				sb.z0 = z0;
				sb.z1 = z0 + per < r1 ? z0 + per : r1;
				sb.off = g->dense_off[l] + z0 * plane;
			}
		}
		const uint32_t dense_total = g->dense_off[kLevels];
		k_grid_dense<<<dim3(sl.count, kGridParts), kGridBlock, 2 * sizeof(float) * kSlabEntries, s>>>(
			a, sl, samples, n, n_dev, g->dgrad_part, dense_total);
		YMX_HIP_CHECK(hipGetLastError());
		LevelOffsets lo;
		for(int l = 0; l <= kLevels; ++l) lo.off[l] = g->dense_off[l];
		lo.mask = g->slab_mask;
		k_grid_reduce<<<(2 * dense_total + 255) / 256, 256, 0, s>>>(g->dgrad_part, dense_total, lo, g->grid_g, g->table);
		YMX_HIP_CHECK(hipGetLastError());
	}
	WgradLayers wl = {{g->dH1, g->dH2, g->dH3, g->dO}, {g->X, g->H1, g->H2, g->H3}};
	k_wgrad<<<dim3(kWgradBlocks, 4), kBlock, 0, s>>>(wl, n, n_dev, g->wgrad_part);
	YMX_HIP_CHECK(hipGetLastError());
	YMX_HIP_CHECK(hipMemsetAsync(g->mlp_g, 0, sizeof(float) * kMlp, s));
	k_wgrad_reduce<<<dim3((4 * 4096 + 255) / 256, kWgradSlices), 256, 0, s>>>(g->wgrad_part, kWgradBlocks, g->mlp_g);
	YMX_HIP_CHECK(hipGetLastError());
	return YMX_OK;
}

int ymx_guide_get_gradients(ymx_guide *g, float *grid, float *mlp, double *loss)
{
	if(!g || !grid || !mlp || !loss)
	{
		ymx_set_error("ymx_guide_get_gradients: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	YMX_HIP_CHECK(hipMemcpy(grid, g->grid_g, sizeof(float) * g->n_grid, hipMemcpyDeviceToHost));
	YMX_HIP_CHECK(hipMemcpy(mlp, g->mlp_g, sizeof(float) * kMlp, hipMemcpyDeviceToHost));
	YMX_HIP_CHECK(hipMemcpy(loss, g->loss, sizeof(double), hipMemcpyDeviceToHost));
	return YMX_OK;
}

int ymx_guide_adam_step(ymx_guide *g, void *stream)
{
	if(!g)
	{
		ymx_set_error("ymx_guide_adam_step: NULL handle");
		return YMX_ERR_INVALID_ARG;
	}
	hipStream_t s = (hipStream_t)stream;
	g->step += 1;
	const float c1 = 1.f - powf(g->p.beta1, (float)g->step);
	const float c2 = 1.f - powf(g->p.beta2, (float)g->step);
	LevelOffsets lo;
	for(int l = 0; l <= kLevels; ++l) lo.off[l] = g->used_off[l];
	lo.mask = (1u << kLevels) - 1u;
	k_adam_grid<<<(2 * g->used_off[kLevels] + 255) / 256, 256, 0, s>>>(g->grid, g->grid_m, g->grid_v, g->grid_g,
	                                                                   g->grid_bf, lo, g->table, g->p.lr, g->p.beta1,
	                                                                   g->p.beta2, g->p.eps, c1, c2);
	YMX_HIP_CHECK(hipGetLastError());
	k_adam<<<(kMlp + 255) / 256, 256, 0, s>>>(g->mlp, g->mlp_m, g->mlp_v, g->mlp_g, nullptr, kMlp, g->p.lr,
	                                         g->p.beta1, g->p.beta2, g->p.eps, c1, c2);
	YMX_HIP_CHECK(hipGetLastError());
	return refresh_bf16(g, s);
}

int ymx_guide_train_step(ymx_guide *g, const ymx_guide_sample *samples, uint32_t n, const uint32_t *n_dev,
                         void *stream)
{
	const int rc = ymx_guide_backward(g, samples, n, n_dev, stream);
	if(rc != YMX_OK) return rc;
	return ymx_guide_adam_step(g, stream);
}

}  // extern "C"
