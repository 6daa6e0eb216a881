// Shared host/device helpers of the HIP graph cut: error handling, owned device
// buffers, wave64 reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "cugip_hip.h"

namespace cgp {

constexpr int32_t kLabelInf = CGP_LABEL_UNREACHABLE;

struct Error : std::runtime_error {
	int code;
	Error(int aCode, const std::string &aMessage) : std::runtime_error(aMessage), code(aCode) {}
};

[[noreturn]] inline void fail(int code, const std::string &message) { throw Error(code, message); }

inline void hip_check(hipError_t err, const char *expr, const char *file, int line)
{
	if (err != hipSuccess) {
		fail(err == hipErrorOutOfMemory ? CGP_ERR_NOMEM : CGP_ERR_DEVICE,
		     std::string(expr) + " failed: " + hipGetErrorString(err) + " (" + file + ":" + std::to_string(line) + ")");
	}
}

#define CGP_HIP(expr) ::cgp::hip_check((expr), #expr, __FILE__, __LINE__)
#define CGP_LAUNCH_CHECK() ::cgp::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// A fixed block of device memory handed out front to back (256-byte aligned pieces), never freed
// piecewise: the memory of the no-alloc entry point. take() fails with CGP_ERR_INVALID when the
// block is exhausted.
struct Arena {
	char *base = nullptr;
	size_t size = 0, used = 0;
	void *take(size_t bytes)
	{
		const size_t start = (used + 255) / 256 * 256;
		if (start > size || bytes > size - start) {
			fail(CGP_ERR_INVALID, "scratch exhausted: " + std::to_string(start + bytes) + " bytes needed, " +
			                          std::to_string(size) + " available");
		}
		used = start + bytes;
		return base + start;
	}
};

// Device allocation (hipMalloc / hipFree), movable, not copyable; or a borrowed piece of memory
// it does not free (an arena piece, a caller's array).
template <typename T>
class DeviceBuffer {
public:
	DeviceBuffer() = default;
	explicit DeviceBuffer(size_t count) { allocate(count); }
	DeviceBuffer(const DeviceBuffer &) = delete;
	DeviceBuffer &operator=(const DeviceBuffer &) = delete;
	DeviceBuffer(DeviceBuffer &&other) noexcept : mPtr(other.mPtr), mCount(other.mCount), mOwned(other.mOwned)
	{
		other.mPtr = nullptr;
		other.mCount = 0;
	}
	DeviceBuffer &operator=(DeviceBuffer &&other) noexcept
	{
		if (this != &other) {
			release();
			mPtr = other.mPtr;
			mCount = other.mCount;
			mOwned = other.mOwned;
			other.mPtr = nullptr;
			other.mCount = 0;
		}
		return *this;
	}
	~DeviceBuffer() { release(); }

	// count elements from `arena` when it is given (no allocation), else from hipMalloc
	void allocate(size_t count, Arena *arena)
	{
		if (!arena) {
			allocate(count);
			return;
		}
		release();
		if (count) {
			borrow(static_cast<T *>(arena->take(count * sizeof(T))), count);
		}
	}
	// use memory owned elsewhere (not freed here)
	void borrow(T *ptr, size_t count)
	{
		release();
		mPtr = ptr;
		mCount = ptr ? count : 0;
		mOwned = false;
	}
	void allocate(size_t count)
	{
		if (count == mCount && mPtr && mOwned) {
			return;
		}
		release();
		if (count == 0) {
			return;
		}
		void *p = nullptr;
		hipError_t err = hipMalloc(&p, count * sizeof(T));
		if (err != hipSuccess) {
			(void)hipGetLastError();
			fail(CGP_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
			                        hipGetErrorString(err));
		}
		mPtr = static_cast<T *>(p);
		mCount = count;
		mOwned = true;
	}
	void release()
	{
		if (mPtr && mOwned) {
			(void)hipFree(mPtr);
		}
		mPtr = nullptr;
		mCount = 0;
		mOwned = true;
	}
	T *get() const { return mPtr; }
	size_t size() const { return mCount; }
	size_t bytes() const { return mCount * sizeof(T); }

private:
	T *mPtr = nullptr;
	size_t mCount = 0;
	bool mOwned = true;
};

// Host staging buffer for device reads: pinned (hipHostMalloc), or with `pageable` plain host
// memory (no HIP allocation; copies into it are slower).
template <typename T>
class PinnedBuffer {
public:
	PinnedBuffer() = default;
	PinnedBuffer(const PinnedBuffer &) = delete;
	PinnedBuffer &operator=(const PinnedBuffer &) = delete;
	~PinnedBuffer() { release(); }
	void allocate(size_t count, bool pageable = false)
	{
		if (count <= mCount && mPtr && pageable == mPageable) {
			return;
		}
		release();
		if (pageable) {
			mHost.assign(count, T());
			mPtr = mHost.data();
		} else {
			void *p = nullptr;
			CGP_HIP(hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault));
			mPtr = static_cast<T *>(p);
		}
		mCount = count;
		mPageable = pageable;
	}
	T *get() const { return mPtr; }
	T &operator[](size_t i) const { return mPtr[i]; }

private:
	void release()
	{
		if (mPtr && !mPageable) {
			(void)hipHostFree(mPtr);
		}
		mHost.clear();
		mPtr = nullptr;
		mCount = 0;
	}
	T *mPtr = nullptr;
	size_t mCount = 0;
	bool mPageable = false;
	std::vector<T> mHost;
};

// ---- device helpers (wave64) ----------------------------------------------------------------

__device__ inline long long wave_sum_i64(long long v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		v += __shfl_xor(v, off, 64);
	}
	return v;
}

__device__ inline int wave_max_i32(int v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		const int o = __shfl_xor(v, off, 64);
		v = o > v ? o : v;
	}
	return v;
}

// Wave maximum and minimum of non-negative values by DPP row shifts and row broadcasts, the result
// collecting in lane 63 and returned wave-uniform; lanes with no source lane take the operation's
// identity (0 for the maximum, all ones for the minimum), so each step is one DPP-combined max/min.
// No LDS round trips, unlike the shuffle version above.
template <bool kMax>
__device__ inline uint32_t wave_reduce_dpp(uint32_t v)
{
	constexpr uint32_t id = kMax ? 0u : 0xffffffffu;
	auto op = [](uint32_t a, uint32_t b) { return kMax ? (a > b ? a : b) : (a < b ? a : b); };
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x111, 0xf, 0xf, false));  // row_shr:1
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x112, 0xf, 0xf, false));  // row_shr:2
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x114, 0xf, 0xf, false));  // row_shr:4
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x118, 0xf, 0xf, false));  // row_shr:8
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));  // row_bcast:15 into rows 1, 3
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xc, 0xf, false));  // row_bcast:31 into rows 2, 3
	return __builtin_amdgcn_readlane(v, 63);
}

// Bitwise OR over the wave (same DPP pattern), wave-uniform.
__device__ inline uint32_t wave_or_dpp(uint32_t v)
{
	v |= __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);
	v |= __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);
	v |= __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);
	v |= __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);
	v |= __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);
	v |= __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);
	return __builtin_amdgcn_readlane(v, 63);
}

__device__ inline int wave_max_nonneg(int v)
{
	return (int)wave_reduce_dpp<true>((uint32_t)v);
}

__device__ inline int wave_min_nonneg(int v)
{
	return (int)wave_reduce_dpp<false>((uint32_t)v);
}

// Block-wide int64 sum, result valid in thread 0. `scratch` holds one slot per wave.
template <int kThreads>
__device__ inline long long block_sum_i64(long long v, long long *scratch)
{
	static_assert(kThreads % 64 == 0, "block size must be a multiple of the wave size");
	v = wave_sum_i64(v);
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	if (lane == 0) {
		scratch[wave] = v;
	}
	__syncthreads();
	long long total = 0;
	if (threadIdx.x == 0) {
#pragma unroll
		for (int w = 0; w < kThreads / 64; ++w) {
			total += scratch[w];
		}
	}
	return total;
}

}  // namespace cgp
