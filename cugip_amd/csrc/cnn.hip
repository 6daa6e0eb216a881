// Learned cut proposal on the matrix cores: 3D convolutions as implicit GEMMs with
// v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32 accumulation), their backward passes, Adam.
//
// One workgroup (4 waves) owns an output brick of 16 x 4 x 4 voxels and stages its input brick
// with a one-voxel halo (18 x 6 x 6 voxels, BF16, one LDS plane per group of 4 channels) in LDS. A
// wave computes one row
// of 16 voxels at a time as D = A x B per tap: A = the 16 x 16 weight block of the tap (held in
// registers for all taps), B = the 16 channels of the 16 voxels shifted by the tap offset (from
// LDS), D = 16 output channels x 16 voxels. In the MFMA's lane layout, lane l holds voxel l % 16
// and channels 4 * (l / 16) .. + 3 of D, so each lane stores 8 contiguous bytes. Inference
// (k_conv_fwd_z) walks several bricks along z per workgroup and keeps the input planes in an LDS
// ring, staging only the planes each brick adds.
#include <cmath>
#include <cstring>

#include "cnn.h"
#include "grid_tile.h"
#include "solver.h"

namespace cgp {

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int BX = 16, BY = 4, BZ = 4;                 // output brick
constexpr int HX = BX + 2, HY = BY + 2, HZ = BZ + 2;   // input brick with halo
constexpr int HV = HX * HY * HZ;
constexpr int BV = BX * BY * BZ;

// bf16 <-> fp32, round to nearest even (no NaN inputs occur here)
__host__ __device__ inline uint16_t f2bf(float f)
{
	uint32_t u;
	__builtin_memcpy(&u, &f, 4);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h)
{
	const uint32_t u = (uint32_t)h << 16;
	float f;
	__builtin_memcpy(&f, &u, 4);
	return f;
}

__host__ __device__ inline int tap_dx(int t) { return t % 3 - 1; }
__host__ __device__ inline int tap_dy(int t) { return (t / 3) % 3 - 1; }
__host__ __device__ inline int tap_dz(int t) { return t / 9 - 1; }

// Parameter offsets in the fp32 vector
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kCnnW1;
__host__ __device__ constexpr int off_w(int layer) { return kOffB1 + kCnnChannels + (layer - 1) * (kCnnW + kCnnChannels); }
__host__ __device__ constexpr int off_b(int layer) { return layer == 0 ? kOffB1 : off_w(layer) + kCnnW; }
constexpr int kOffHeadW = off_w(4);  // after layer 3 (index 3 = fourth conv): head follows
constexpr int kOffHeadB = kOffHeadW + kCnnChannels;
static_assert(kOffHeadB + 1 == kCnnParams, "parameter layout");

// Packed A fragments (bf16): forward layer 0 has 7 groups of 4 taps, layers 1-3 27 taps each.
constexpr int kPackL0 = 7 * 64 * 4;
constexpr int kPackL = 27 * 64 * 4;
__host__ __device__ constexpr int pack_fwd_off(int layer) { return layer == 0 ? 0 : kPackL0 + (layer - 1) * kPackL; }
constexpr int kPackFwd = kPackL0 + 3 * kPackL;
constexpr int kPackBwd = 3 * kPackL;  // layers 1-3 (the input layer needs no data gradient)

// Fills the MFMA A operands from the fp32 parameters. Forward (A rows = output channel, k = input
// channel): lane l holds W[t][4*(l/16)+j][l%16]. Layer 0 packs taps 4*g .. 4*g+3 into one K = 16
// block: lane l holds W1[4*g + l/16][j][l%16] (zero past tap 26). Backward data (A rows = input
// channel, k = output channel): lane l holds W[t][l%16][4*(l/16)+j].
__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ p, uint16_t *__restrict__ fwd,
                                             uint16_t *__restrict__ bwd)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < kPackFwd) {
		const int j = i & 3, lane = (i >> 2) & 63, m = i >> 8;
		const int col = lane & 15, grp = lane >> 4;
		float w = 0.0f;
		if (i < kPackL0) {
			const int t = 4 * m + grp;
			if (t < kCnnTaps) {
				w = p[kOffW1 + (t * 4 + j) * kCnnChannels + col];
			}
		} else {
			const int layer = 1 + (i - kPackL0) / kPackL;
			const int t = ((i - kPackL0) % kPackL) >> 8;
			w = p[off_w(layer) + (t * kCnnChannels + 4 * grp + j) * kCnnChannels + col];
		}
		fwd[i] = f2bf(w);
	}
	if (i < kPackBwd) {
		const int j = i & 3, lane = (i >> 2) & 63;
		const int layer = 1 + i / kPackL;
		const int t = (i % kPackL) >> 8;
		const int col = lane & 15, grp = lane >> 4;
		bwd[i] = f2bf(p[off_w(layer) + (t * kCnnChannels + col) * kCnnChannels + 4 * grp + j]);
	}
}

// Network input: intensity / 255, source and sink t-link / 1000, 0 (four bf16 channels).
__global__ __launch_bounds__(256) void k_cnn_input(int64_t n, const uint8_t *__restrict__ vol,
                                                  const int32_t *__restrict__ cs, const int32_t *__restrict__ ct,
                                                  uint16_t *__restrict__ x0)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
		ushort4 q;
		q.x = f2bf((float)vol[v] * (1.0f / 255.0f));
		q.y = f2bf((float)cs[v] * 1e-3f);
		q.z = f2bf((float)ct[v] * 1e-3f);
		q.w = 0;
		*reinterpret_cast<ushort4 *>(x0 + 4 * v) = q;
	}
}

// Stage the input brick (channel-last, CIN channels) with its halo; zero outside the volume.
template <int CIN>
__device__ inline void stage_brick(uint16_t *lds, const uint16_t *__restrict__ in, int nx, int ny, int nz, int bx,
                                   int by, int bz)
{
	constexpr int kParts = CIN / 4;  // 8-byte units per voxel
	for (int e = threadIdx.x; e < HV * kParts; e += blockDim.x) {
		const int hv = e / kParts, part = e % kParts;
		const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
		const int x = bx + hx - 1, y = by + hy - 1, z = bz + hz - 1;
		uint2 val = make_uint2(0u, 0u);
		if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
			const int64_t v = ((int64_t)z * ny + y) * nx + x;
			val = *reinterpret_cast<const uint2 *>(in + v * CIN + part * 4);
		}
		*reinterpret_cast<uint2 *>(lds + hv * CIN + part * 4) = val;
	}
}

// Stage the input brick with its halo as one plane per group of 4 channels (8 bytes a voxel), so
// that the 16 lanes of an MFMA B fragment, which read one channel group of 16 consecutive voxels,
// read 128 contiguous bytes (channel-last, their 8-byte reads would stride 32 bytes and collide in
// the LDS banks). Zero outside the volume.
template <int CIN>
__device__ inline void stage_brick_planar(uint16_t *lds, const uint16_t *__restrict__ in, int nx, int ny, int nz,
                                          int bx, int by, int bz)
{
	constexpr int kParts = CIN / 4;
	for (int e = threadIdx.x; e < HV * kParts; e += blockDim.x) {
		const int hv = e / kParts, part = e % kParts;
		const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
		const int x = bx + hx - 1, y = by + hy - 1, z = bz + hz - 1;
		uint2 val = make_uint2(0u, 0u);
		if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
			const int64_t v = ((int64_t)z * ny + y) * nx + x;
			val = *reinterpret_cast<const uint2 *>(in + v * CIN + part * 4);
		}
		*reinterpret_cast<uint2 *>(lds + (part * HV + hv) * 4) = val;
	}
}

// Convolution layer. Forward (BWD false): out = ReLU(conv(in, W) + b) in bf16; with HEAD, also
// logit = w_head . out + b_head (the sigmoid head fused into the last layer). Backward data (BWD
// true, CIN 16): out = conv^T(g, W) masked by (mask_act > 0), the gradient at the pre-activation
// of the layer below.
template <int CIN, bool BWD, bool HEAD>
__global__ __launch_bounds__(256) void k_conv(int nx, int ny, int nz, const uint16_t *__restrict__ in,
                                             const uint16_t *__restrict__ pack, const float *__restrict__ bias,
                                             const uint16_t *__restrict__ mask_act, uint16_t *__restrict__ out,
                                             const float *__restrict__ head, float *__restrict__ logits)
{
	__shared__ __attribute__((aligned(16))) uint16_t lds[HV * CIN];
	const int bx = blockIdx.x * BX, by = blockIdx.y * BY, bz = blockIdx.z * BZ;
	stage_brick_planar<CIN>(lds, in, nx, ny, nz, bx, by, bz);
	constexpr int NM = CIN == 16 ? 27 : 7;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int col = lane & 15, grp = lane >> 4;
	short4_t a[NM];
#pragma unroll
	for (int m = 0; m < NM; ++m) {
		a[m] = *reinterpret_cast<const short4_t *>(pack + (m * 64 + lane) * 4);
	}
	float bias4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
	float hw4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
	if (!BWD) {
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			bias4[j] = bias[4 * grp + j];
			if (HEAD) {
				hw4[j] = head[4 * grp + j];
			}
		}
	}
	__syncthreads();
#pragma unroll 1
	for (int i = 0; i < 4; ++i) {
		const int row = wave * 4 + i;
		const int ly = row % BY, lz = row / BY;
		float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
		for (int m = 0; m < NM; ++m) {
			int t;
			int c0;
			if (CIN == 16) {
				t = BWD ? kCnnTaps - 1 - m : m;  // backward reads g at u - off(t) = u + off(26 - t)
				c0 = 4 * grp;
			} else {
				t = 4 * m + grp;
				t = t < kCnnTaps ? t : kCnnTaps - 1;  // padded taps have zero weights
				c0 = 0;
			}
			const int hv = ((lz + 1 + tap_dz(t)) * HY + (ly + 1 + tap_dy(t))) * HX + (col + 1 + tap_dx(t));
			const short4_t b = *reinterpret_cast<const short4_t *>(lds + ((c0 / 4) * HV + hv) * 4);
			acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b, acc, 0, 0, 0);
		}
		const int x = bx + col, y = by + ly, z = bz + lz;
		const bool valid = x < nx && y < ny && z < nz;
		const int64_t v = ((int64_t)z * ny + y) * nx + x;
		ushort4 q;
		uint16_t r[4];
		float head_part = 0.0f;
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			float val = acc[j];
			if (!BWD) {
				val += bias4[j];
				val = val > 0.0f ? val : 0.0f;
				r[j] = f2bf(val);
				if (HEAD) {
					head_part += hw4[j] * bf2f(r[j]);
				}
			} else {
				r[j] = 0;
			}
		}
		if (BWD && valid) {
			const ushort4 mk = *reinterpret_cast<const ushort4 *>(mask_act + v * 16 + 4 * grp);
			r[0] = mk.x ? f2bf(acc[0]) : 0;
			r[1] = mk.y ? f2bf(acc[1]) : 0;
			r[2] = mk.z ? f2bf(acc[2]) : 0;
			r[3] = mk.w ? f2bf(acc[3]) : 0;
		}
		q.x = r[0];
		q.y = r[1];
		q.z = r[2];
		q.w = r[3];
		if (HEAD) {
			head_part += __shfl_xor(head_part, 16, 64);
			head_part += __shfl_xor(head_part, 32, 64);
		}
		if (valid) {
			*reinterpret_cast<ushort4 *>(out + v * 16 + 4 * grp) = q;
			if (HEAD && grp == 0) {
				logits[v] = head_part + head[kCnnChannels];
			}
		}
	}
}

// Forward layer for inference, marching along z: workgroup (bx, by, bz) computes kZBricks
// consecutive 16 x 4 x 4 output bricks in z. The input planes stay in an LDS ring of 8 z-planes
// (one plane per 4-channel group as in k_conv), so each brick stages only the 4 planes it adds
// instead of its whole 6-plane halo box. Same arithmetic and output as k_conv<CIN, false, HEAD>.
constexpr int kZBricks = 8;
constexpr int kZSlots = 8;

// Per-thread staging map of one input z-plane of the brick box (18 x 6 voxels, CIN / 4 channel
// groups): thread t stages entries t, t + 256, ...; their slice offsets in global memory and their
// LDS offsets (without the ring slot) are fixed for the workgroup, so staging a plane is a load and
// a store per entry.
template <int CIN>
struct PlaneMap {
	static constexpr int kParts = CIN / 4;
	static constexpr int PL = HX * HY;
	static constexpr int kPerPlane = PL * kParts;
	static constexpr int kIters = (kPerPlane + 255) / 256;
	uint32_t goff[kIters];  // uint16 offset of the entry in a z-slice of `in`
	uint32_t loff[kIters];  // uint16 offset of the entry in LDS, ring slot 0
	uint32_t inside = 0;    // bit i: entry i exists and its (x, y) lies in the volume
	uint32_t exists = 0;    // bit i: entry i exists

	__device__ void init(int nx, int ny, int bx, int by)
	{
#pragma unroll
		for (int i = 0; i < kIters; ++i) {
			const int e = (int)threadIdx.x + i * 256;
			const int part = e % kParts, pv = e / kParts;
			const int hx = pv % HX, hy = pv / HX;
			const int x = bx + hx - 1, y = by + hy - 1;
			const bool ex = e < kPerPlane;
			exists |= ex ? (1u << i) : 0u;
			inside |= (ex && x >= 0 && x < nx && y >= 0 && y < ny) ? (1u << i) : 0u;
			goff[i] = (inside >> i) & 1u ? (uint32_t)((y * nx + x) * CIN + part * 4) : 0u;
			loff[i] = (uint32_t)((part * kZSlots * PL + pv) * 4);
		}
	}

	__device__ void stage(uint16_t *lds, const uint16_t *__restrict__ in, int nx, int ny, int nz, int z) const
	{
		const bool zin = z >= 0 && z < nz;
		const uint16_t *slice = in + (zin ? (int64_t)z * nx * ny * CIN : 0);
		const uint32_t soff = ((uint32_t)(z + kZSlots) & (uint32_t)(kZSlots - 1)) * PL * 4;
#pragma unroll
		for (int i = 0; i < kIters; ++i) {
			if ((exists >> i) & 1u) {
				uint2 val = make_uint2(0u, 0u);
				if (zin && ((inside >> i) & 1u)) {
					val = *reinterpret_cast<const uint2 *>(slice + goff[i]);
				}
				*reinterpret_cast<uint2 *>(lds + loff[i] + soff) = val;
			}
		}
	}
};

template <int CIN, bool HEAD>
__global__ __launch_bounds__(256) void k_conv_fwd_z(int nx, int ny, int nz, const uint16_t *__restrict__ in,
                                                   const uint16_t *__restrict__ pack, const float *__restrict__ bias,
                                                   uint16_t *__restrict__ out, const float *__restrict__ head,
                                                   float *__restrict__ logits)
{
	static_assert(kZSlots == 8, "ring slots are taken modulo 8");
	constexpr int PL = HX * HY;
	__shared__ __attribute__((aligned(16))) uint16_t lds[kZSlots * PL * CIN];
	const int bx = blockIdx.x * BX, by = blockIdx.y * BY, bz0 = blockIdx.z * (kZBricks * BZ);
	constexpr int NM = CIN == 16 ? 27 : 7;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int col = lane & 15, grp = lane >> 4;
	short4_t a[NM];
#pragma unroll
	for (int m = 0; m < NM; ++m) {
		a[m] = *reinterpret_cast<const short4_t *>(pack + (m * 64 + lane) * 4);
	}
	float bias4[4], hw4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
	for (int j = 0; j < 4; ++j) {
		bias4[j] = bias[4 * grp + j];
		if (HEAD) {
			hw4[j] = head[4 * grp + j];
		}
	}
	PlaneMap<CIN> map;
	map.init(nx, ny, bx, by);
#pragma unroll 1
	for (int p = -1; p <= BZ; ++p) {
		map.stage(lds, in, nx, ny, nz, bz0 + p);
	}
	// the lane's LDS base: its channel group's planes, its voxel column (x + 1 for the halo)
	const uint32_t lbase = (uint32_t)((((CIN == 16 ? grp : 0) * kZSlots * PL) + col + 1) * 4);
	const int lz = wave;
	for (int zb = 0; zb < kZBricks; ++zb) {
		const int bz = bz0 + zb * BZ;
		if (bz >= nz) {
			break;  // uniform across the workgroup
		}
		if (zb > 0) {
			__syncthreads();  // the previous brick's reads of the slots about to be refilled are done
#pragma unroll 1
			for (int p = 1; p <= BZ; ++p) {
				map.stage(lds, in, nx, ny, nz, bz + p);
			}
		}
		__syncthreads();
		// ring slot offsets of the planes z - 1, z, z + 1 of this wave's output slice
		uint32_t sdz[3];
#pragma unroll
		for (int d = 0; d < 3; ++d) {
			sdz[d] = ((uint32_t)(bz + lz + d - 1 + kZSlots) & (uint32_t)(kZSlots - 1)) * PL * 4;
		}
#pragma unroll 1
		for (int ly = 0; ly < BY; ++ly) {
			float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
			for (int m = 0; m < NM; ++m) {
				short4_t b;
				if constexpr (CIN == 16) {
					const int t = m;  // compile-time offsets: the reads take immediate offsets
					const uint32_t off = sdz[tap_dz(t) + 1] +
					                     (uint32_t)((((ly + 1 + tap_dy(t)) * HX) + tap_dx(t)) * 4);
					b = *reinterpret_cast<const short4_t *>(lds + lbase + off);
				} else {
					int t = 4 * m + grp;
					t = t < kCnnTaps ? t : kCnnTaps - 1;  // padded taps have zero weights
					const uint32_t off = sdz[tap_dz(t) + 1] +
					                     (uint32_t)((((ly + 1 + tap_dy(t)) * HX) + tap_dx(t)) * 4);
					b = *reinterpret_cast<const short4_t *>(lds + lbase + off);
				}
				acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b, acc, 0, 0, 0);
			}
			const int x = bx + col, y = by + ly, z = bz + lz;
			const bool valid = x < nx && y < ny && z < nz;
			const int64_t v = ((int64_t)z * ny + y) * nx + x;
			ushort4 q;
			uint16_t r[4];
			float head_part = 0.0f;
#pragma unroll
			for (int j = 0; j < 4; ++j) {
				float val = acc[j] + bias4[j];
				val = val > 0.0f ? val : 0.0f;
				r[j] = f2bf(val);
				if (HEAD) {
					head_part += hw4[j] * bf2f(r[j]);
				}
			}
			q.x = r[0];
			q.y = r[1];
			q.z = r[2];
			q.w = r[3];
			if (HEAD) {
				head_part += __shfl_xor(head_part, 16, 64);
				head_part += __shfl_xor(head_part, 32, 64);
			}
			if (valid) {
				*reinterpret_cast<ushort4 *>(out + v * 16 + 4 * grp) = q;
				if (HEAD && grp == 0) {
					logits[v] = head_part + head[kCnnChannels];
				}
			}
		}
	}
}

// Weight gradient of one layer: dW[t][ci][co] = sum_v in[v + off(t)][ci] * g[v][co] and
// db[co] = sum_v g[v][co], as D = A x B over chunks of 16 voxels: A = in^T (ci x voxel, from the
// input brick), B = g (voxel x co). Wave w owns taps w, w + 4, ...; a workgroup walks bricks with a
// grid stride and writes its partial sums (27*CIN*16 weights, then 16 biases) to partial[block].
template <int CIN>
__global__ __launch_bounds__(256) void k_conv_wgrad(int nx, int ny, int nz, int nbx, int nby, int nbz,
                                                   const uint16_t *__restrict__ in, const uint16_t *__restrict__ g,
                                                   float *__restrict__ partial)
{
	__shared__ __attribute__((aligned(16))) uint16_t lin[HV * CIN];
	__shared__ __attribute__((aligned(16))) uint16_t lg[BV * 16];
	__shared__ float dbs[256];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int col = lane & 15, grp = lane >> 4;
	constexpr int kMyTaps = 7;  // ceil(27 / 4)
	float4_t acc[kMyTaps];
#pragma unroll
	for (int k = 0; k < kMyTaps; ++k) {
		acc[k] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
	}
	float db = 0.0f;
	const int nbricks = nbx * nby * nbz;
	for (int brick = blockIdx.x; brick < nbricks; brick += gridDim.x) {
		const int bx = (brick % nbx) * BX, by = ((brick / nbx) % nby) * BY, bz = (brick / (nbx * nby)) * BZ;
		__syncthreads();  // the previous brick's LDS reads are done
		stage_brick<CIN>(lin, in, nx, ny, nz, bx, by, bz);
		for (int e = threadIdx.x; e < BV * 4; e += blockDim.x) {
			const int bv = e >> 2, part = e & 3;
			const int lx = bv % BX, ly = (bv / BX) % BY, lz = bv / (BX * BY);
			const int x = bx + lx, y = by + ly, z = bz + lz;
			uint2 val = make_uint2(0u, 0u);
			if (x < nx && y < ny && z < nz) {
				const int64_t v = ((int64_t)z * ny + y) * nx + x;
				val = *reinterpret_cast<const uint2 *>(g + v * 16 + part * 4);
			}
			*reinterpret_cast<uint2 *>(lg + bv * 16 + part * 4) = val;
		}
		__syncthreads();
		// bias: thread (v16, co) sums 16 voxels of channel co
		{
			const int co = threadIdx.x & 15, v16 = threadIdx.x >> 4;
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				db += bf2f(lg[(v16 * 16 + k) * 16 + co]);
			}
		}
#pragma unroll
		for (int k = 0; k < kMyTaps; ++k) {
			const int t = wave + 4 * k;
			if (t >= kCnnTaps) {
				continue;
			}
#pragma unroll 1
			for (int row = 0; row < BY * BZ; ++row) {
				const int ly = row % BY, lz = row / BY;
				short4_t av, bv;
#pragma unroll
				for (int j = 0; j < 4; ++j) {
					const int x = 4 * grp + j;  // voxel of the chunk (k index)
					const int hv = ((lz + 1 + tap_dz(t)) * HY + (ly + 1 + tap_dy(t))) * HX + (x + 1 + tap_dx(t));
					av[j] = col < CIN ? (short)lin[hv * CIN + (col < CIN ? col : 0)] : (short)0;
					bv[j] = (short)lg[(row * BX + x) * 16 + col];
				}
				acc[k] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[k], 0, 0, 0);
			}
		}
	}
	float *mine = partial + (int64_t)blockIdx.x * (kCnnTaps * CIN * 16 + 16);
#pragma unroll
	for (int k = 0; k < kMyTaps; ++k) {
		const int t = wave + 4 * k;
		if (t >= kCnnTaps) {
			continue;
		}
		// D[ci][co]: lane holds co = col, ci = 4 * grp + j
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const int ci = 4 * grp + j;
			if (ci < CIN) {
				mine[(t * CIN + ci) * 16 + col] = acc[k][j];
			}
		}
	}
	dbs[threadIdx.x] = db;
	__syncthreads();
	if (threadIdx.x < 16) {
		float s = 0.0f;
		for (int k = threadIdx.x; k < 256; k += 16) {
			s += dbs[k];
		}
		mine[kCnnTaps * CIN * 16 + threadIdx.x] = s;
	}
}

// out[i] = sum over blocks of partial[block][i]
__global__ __launch_bounds__(256) void k_reduce_partials(const float *__restrict__ partial, int nblocks, int nvals,
                                                        float *__restrict__ out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nvals) {
		return;
	}
	float s = 0.0f;
	for (int b = 0; b < nblocks; ++b) {
		s += partial[(int64_t)b * nvals + i];
	}
	out[i] = s;
}

// Head backward: dz = (sigmoid(z) - y) / n; the gradient at layer 4's pre-activation is
// dz * w_head where layer 4's output is positive. Per-block partials: dw_head[16], then
// (double accumulators) loss and db_head.
__global__ __launch_bounds__(256) void k_head_bwd(int64_t n, const float *__restrict__ logits,
                                                 const uint8_t *__restrict__ label, const uint16_t *__restrict__ act,
                                                 const float *__restrict__ head, uint16_t *__restrict__ g,
                                                 float *__restrict__ partial, double *__restrict__ scalars)
{
	__shared__ float red[16][257];
	__shared__ double rl[256], rb[256];
	const float inv_n = 1.0f / (float)n;
	float gw[16];
#pragma unroll
	for (int c = 0; c < 16; ++c) {
		gw[c] = 0.0f;
	}
	double loss = 0.0, dbh = 0.0;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
		const float z = logits[v];
		const float y = label[v] ? 1.0f : 0.0f;
		const float p = 1.0f / (1.0f + expf(-z));
		const float dz = (p - y) * inv_n;
		loss += (double)(fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z))) - y * z);
		dbh += (double)dz;
		uint16_t out[16];
#pragma unroll
		for (int c = 0; c < 16; ++c) {
			const float a = bf2f(act[v * 16 + c]);
			gw[c] += dz * a;
			out[c] = a > 0.0f ? f2bf(dz * head[c]) : 0;
		}
#pragma unroll
		for (int c = 0; c < 16; c += 4) {
			ushort4 q;
			q.x = out[c];
			q.y = out[c + 1];
			q.z = out[c + 2];
			q.w = out[c + 3];
			*reinterpret_cast<ushort4 *>(g + v * 16 + c) = q;
		}
	}
#pragma unroll
	for (int c = 0; c < 16; ++c) {
		red[c][threadIdx.x] = gw[c];
	}
	rl[threadIdx.x] = loss;
	rb[threadIdx.x] = dbh;
	__syncthreads();
	if (threadIdx.x < 16) {
		float s = 0.0f;
		for (int k = 0; k < 256; ++k) {
			s += red[threadIdx.x][k];
		}
		partial[(int64_t)blockIdx.x * 16 + threadIdx.x] = s;
	}
	if (threadIdx.x == 0) {
		double a = 0.0, b = 0.0;
		for (int k = 0; k < 256; ++k) {
			a += rl[k];
			b += rb[k];
		}
		atomicAdd(&scalars[0], a);
		atomicAdd(&scalars[1], b);
	}
}

// Adam on the fp32 parameters; the padding input channel of layer 0 stays zero.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g,
                                             float *__restrict__ m, float *__restrict__ v, float lr, float c1,
                                             float c2)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= kCnnParams) {
		return;
	}
	if (i < kCnnW1 && ((i / kCnnChannels) & 3) == 3) {
		p[i] = 0.0f;
		return;
	}
	const float gi = g[i];
	const float mi = 0.9f * m[i] + 0.1f * gi;
	const float vi = 0.999f * v[i] + 0.001f * gi * gi;
	m[i] = mi;
	v[i] = vi;
	p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

__global__ __launch_bounds__(256) void k_threshold(int64_t n, const float *__restrict__ logits,
                                                  uint8_t *__restrict__ side)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
		side[v] = logits[v] > 0.0f ? 1 : 0;
	}
}

__global__ __launch_bounds__(256) void k_sigmoid(int64_t n, const float *__restrict__ logits, float *__restrict__ prob)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
		prob[v] = 1.0f / (1.0f + expf(-logits[v]));
	}
}

int blocks_of(int64_t n, int cap = 8192)
{
	int64_t b = (n + 255) / 256;
	return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

CnnModel::CnnModel(hipStream_t stream) : mStream(stream), mHost(kCnnParams, 0.0f)
{
	mParams.allocate(kCnnParams);
	mGrad.allocate(kCnnParams);
	mAdamM.allocate(kCnnParams);
	mAdamV.allocate(kCnnParams);
	mPackFwd.allocate(kPackFwd);
	mPackBwd.allocate(kPackBwd);
	mScalars.allocate(2);
	CGP_HIP(hipMemsetAsync(mParams.get(), 0, kCnnParams * sizeof(float), mStream));
	CGP_HIP(hipMemsetAsync(mAdamM.get(), 0, kCnnParams * sizeof(float), mStream));
	CGP_HIP(hipMemsetAsync(mAdamV.get(), 0, kCnnParams * sizeof(float), mStream));
	upload_packed();
}

void CnnModel::set_params(const float *params)
{
	for (int i = 0; i < kCnnParams; ++i) {
		if (!std::isfinite(params[i])) {
			fail(CGP_ERR_INVALID, "parameter " + std::to_string(i) + " is not finite");
		}
	}
	std::memcpy(mHost.data(), params, kCnnParams * sizeof(float));
	for (int t = 0; t < kCnnTaps; ++t) {
		for (int co = 0; co < kCnnChannels; ++co) {
			mHost[(size_t)(kOffW1 + (t * 4 + 3) * kCnnChannels + co)] = 0.0f;  // padding channel
		}
	}
	CGP_HIP(hipMemcpyAsync(mParams.get(), mHost.data(), kCnnParams * sizeof(float), hipMemcpyHostToDevice, mStream));
	CGP_HIP(hipMemsetAsync(mAdamM.get(), 0, kCnnParams * sizeof(float), mStream));
	CGP_HIP(hipMemsetAsync(mAdamV.get(), 0, kCnnParams * sizeof(float), mStream));
	mStep = 0;
	upload_packed();
}

void CnnModel::get_params(float *params)
{
	CGP_HIP(hipMemcpyAsync(params, mParams.get(), kCnnParams * sizeof(float), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

void CnnModel::upload_packed()
{
	hipLaunchKernelGGL(k_pack, dim3((kPackFwd + 255) / 256), dim3(256), 0, mStream, mParams.get(), mPackFwd.get(),
	                   mPackBwd.get());
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipStreamSynchronize(mStream));
}

void CnnModel::ensure(int64_t vox)
{
	if (vox == mVox) {
		return;
	}
	mX0.allocate((size_t)vox * 4);
	for (auto &a : mAct) {
		a.allocate((size_t)vox * 16);
	}
	for (auto &g : mG) {
		g.allocate((size_t)vox * 16);
	}
	mLogit.allocate(vox);
	mVolume.allocate(vox);
	mSide.allocate(vox);
	mVox = vox;
}

void CnnModel::forward(GridSolver &grid, const uint8_t *volume, bool device_volume, bool)
{
	if (!volume) {
		volume = grid.device_volume();  // the grid was built from its volume on the device
		device_volume = true;
		if (!volume) {
			fail(CGP_ERR_INVALID, "no volume: pass one, or build the grid with cgp_set_grid_volume");
		}
	}
	const GridSolver::DeviceGrid d = grid.device_grid(0);
	ensure((int64_t)d.nx * d.ny * d.nz);
	mNx = d.nx;
	mNy = d.ny;
	mNz = d.nz;
	const uint8_t *vol = volume;
	if (!device_volume) {
		CGP_HIP(hipMemcpyAsync(mVolume.get(), volume, mVox, hipMemcpyHostToDevice, mStream));
		vol = mVolume.get();
	}
	hipLaunchKernelGGL(k_cnn_input, dim3(blocks_of(mVox)), dim3(256), 0, mStream, mVox, vol, d.cs, d.ct, mX0.get());
	// inference layers march along z, kZBricks bricks a workgroup
	const dim3 gridz((mNx + BX - 1) / BX, (mNy + BY - 1) / BY, (mNz + BZ * kZBricks - 1) / (BZ * kZBricks));
	const float *p = mParams.get();
	hipLaunchKernelGGL((k_conv_fwd_z<4, false>), gridz, dim3(256), 0, mStream, mNx, mNy, mNz, mX0.get(),
	                   mPackFwd.get() + pack_fwd_off(0), p + off_b(0), mAct[0].get(), (const float *)nullptr,
	                   (float *)nullptr);
	for (int layer = 1; layer <= 2; ++layer) {
		hipLaunchKernelGGL((k_conv_fwd_z<16, false>), gridz, dim3(256), 0, mStream, mNx, mNy, mNz,
		                   mAct[layer - 1].get(), mPackFwd.get() + pack_fwd_off(layer), p + off_b(layer),
		                   mAct[layer].get(), (const float *)nullptr, (float *)nullptr);
	}
	hipLaunchKernelGGL((k_conv_fwd_z<16, true>), gridz, dim3(256), 0, mStream, mNx, mNy, mNz, mAct[2].get(),
	                   mPackFwd.get() + pack_fwd_off(3), p + off_b(3), mAct[3].get(), p + kOffHeadW, mLogit.get());
	hipLaunchKernelGGL(k_threshold, dim3(blocks_of(mVox)), dim3(256), 0, mStream, mVox, mLogit.get(), mSide.get());
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipStreamSynchronize(mStream));
}

const uint8_t *CnnModel::predicted_side()
{
	if (!mVox) {
		fail(CGP_ERR_STATE, "no prediction yet");
	}
	return mSide.get();
}

void CnnModel::prediction_to_host(uint8_t *side, float *prob)
{
	if (side) {
		CGP_HIP(hipMemcpyAsync(side, mSide.get(), mVox, hipMemcpyDeviceToHost, mStream));
	}
	if (prob) {
		DeviceBuffer<float> p(mVox);
		hipLaunchKernelGGL(k_sigmoid, dim3(blocks_of(mVox)), dim3(256), 0, mStream, mVox, mLogit.get(), p.get());
		CGP_LAUNCH_CHECK();
		CGP_HIP(hipMemcpyAsync(prob, p.get(), mVox * sizeof(float), hipMemcpyDeviceToHost, mStream));
		CGP_HIP(hipStreamSynchronize(mStream));
	}
	CGP_HIP(hipStreamSynchronize(mStream));
}

void CnnModel::run_backward(GridSolver &grid, double &loss)
{
	const uint8_t *label = grid.device_source_side(0);
	const float *p = mParams.get();
	CGP_HIP(hipMemsetAsync(mScalars.get(), 0, 2 * sizeof(double), mStream));
	const int head_blocks = blocks_of(mVox, 2048);
	const int nbx = (mNx + BX - 1) / BX, nby = (mNy + BY - 1) / BY, nbz = (mNz + BZ - 1) / BZ;
	const int wg_blocks = std::min(nbx * nby * nbz, 2048);
	const int per16 = kCnnTaps * 16 * 16 + 16, per4 = kCnnTaps * 4 * 16 + 16;
	mPartials.allocate((size_t)std::max(head_blocks * 16, wg_blocks * per16));
	// head
	hipLaunchKernelGGL(k_head_bwd, dim3(head_blocks), dim3(256), 0, mStream, mVox, mLogit.get(), label, mAct[3].get(),
	                   p + kOffHeadW, mG[0].get(), mPartials.get(), mScalars.get());
	hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, mStream, mPartials.get(), head_blocks, 16,
	                   mGrad.get() + kOffHeadW);
	const dim3 grid3(nbx, nby, nbz);
	int cur = 0;
	for (int layer = 3; layer >= 1; --layer) {
		// weight gradient of `layer` (input: activation of layer - 1, gradient at its pre-activation in mG[cur])
		hipLaunchKernelGGL((k_conv_wgrad<16>), dim3(wg_blocks), dim3(256), 0, mStream, mNx, mNy, mNz, nbx, nby, nbz,
		                   mAct[layer - 1].get(), mG[cur].get(), mPartials.get());
		hipLaunchKernelGGL(k_reduce_partials, dim3((per16 + 255) / 256), dim3(256), 0, mStream, mPartials.get(),
		                   wg_blocks, per16, mGrad.get() + off_w(layer));
		// gradient at the pre-activation of layer - 1
		hipLaunchKernelGGL((k_conv<16, true, false>), grid3, dim3(256), 0, mStream, mNx, mNy, mNz, mG[cur].get(),
		                   mPackBwd.get() + (layer - 1) * kPackL, (const float *)nullptr, mAct[layer - 1].get(),
		                   mG[1 - cur].get(), (const float *)nullptr, (float *)nullptr);
		cur = 1 - cur;
	}
	hipLaunchKernelGGL((k_conv_wgrad<4>), dim3(wg_blocks), dim3(256), 0, mStream, mNx, mNy, mNz, nbx, nby, nbz,
	                   mX0.get(), mG[cur].get(), mPartials.get());
	hipLaunchKernelGGL(k_reduce_partials, dim3((per4 + 255) / 256), dim3(256), 0, mStream, mPartials.get(), wg_blocks,
	                   per4, mGrad.get() + kOffW1);
	CGP_LAUNCH_CHECK();
	double host[2];
	CGP_HIP(hipMemcpyAsync(host, mScalars.get(), sizeof(host), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
	const float dbh = (float)host[1];
	CGP_HIP(hipMemcpyAsync(mGrad.get() + kOffHeadB, &dbh, sizeof(float), hipMemcpyHostToDevice, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
	loss = host[0] / (double)mVox;
}

double CnnModel::gradients(GridSolver &grid, const uint8_t *volume, bool device_volume, float *grad)
{
	forward(grid, volume, device_volume, true);
	double loss = 0.0;
	run_backward(grid, loss);
	if (grad) {
		CGP_HIP(hipMemcpyAsync(grad, mGrad.get(), kCnnParams * sizeof(float), hipMemcpyDeviceToHost, mStream));
		CGP_HIP(hipStreamSynchronize(mStream));
		for (int t = 0; t < kCnnTaps; ++t) {
			for (int co = 0; co < kCnnChannels; ++co) {
				grad[kOffW1 + (t * 4 + 3) * kCnnChannels + co] = 0.0f;  // padding channel
			}
		}
	}
	return loss;
}

double CnnModel::train_step(GridSolver &grid, const uint8_t *volume, bool device_volume, float learning_rate)
{
	if (!(learning_rate > 0.0f) || !std::isfinite(learning_rate)) {
		fail(CGP_ERR_INVALID, "learning rate must be positive and finite");
	}
	const double loss = gradients(grid, volume, device_volume, nullptr);
	++mStep;
	const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)mStep));
	const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)mStep));
	hipLaunchKernelGGL(k_adam, dim3((kCnnParams + 255) / 256), dim3(256), 0, mStream, mParams.get(), mGrad.get(),
	                   mAdamM.get(), mAdamV.get(), learning_rate, c1, c2);
	CGP_LAUNCH_CHECK();
	upload_packed();
	return loss;
}

}  // namespace cgp
