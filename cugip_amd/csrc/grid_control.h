// Device-driven solve of the grid path: the relabel / discharge / closure loop of the reference
// (MinCut::compute, include/cugip/advanced_operations/detail/graph_cut_implementation.hpp:122-198)
// with its decisions taken on the device. A control block in device memory holds the stage the
// solve is in; every kernel of a stage is gated by it (a compaction whose gate is closed lists no
// tile, so the passes that take its list do nothing), and single-workgroup controller kernels at
// the end of each stage move the solve on. The host captures init + one round of stages in a HIP
// graph once and replays it, reading the control block once per replay, until the solve is done.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cgp {

// Stages of a device-driven solve (SolveControl::mode).
enum SolveMode : int32_t {
	kModeInit = 0,          // init kernel pending
	kModeRelabelNew = 1,    // global relabel from scratch (labels reset, every tile visited once)
	kModeRelabelCont = 2,   // global relabel continued: dirty tiles left at the last check
	kModeDischarge = 3,     // discharge phases
	kModeClosureNew = 4,    // termination test: closure of the excess, stops when it reaches t
	kModeClosureCont = 5,
	kModeFinalNew = 6,      // S_min closure after a relabel found nothing active (labels exact)
	kModeFinalCont = 7,
	kModeDoneClosed = 8,    // done: the closure test found no way to t (S_min = the closure)
	kModeDoneExact = 9,     // done: the last relabel found nothing active (labels exact, S_min closed)
	kModeFailed = 10,       // invalid capacities (status bits 0-1 in SolveControl::status)
	kModeNeedWide = 11,     // residuals do not fit int16: the discharge must stage them as int32
	kModeNotConverged = 12  // the round cap was reached
};

constexpr unsigned mode_bit(int m) { return 1u << m; }
constexpr unsigned kGateRelabelNew = mode_bit(kModeRelabelNew);
constexpr unsigned kGateRelabel = mode_bit(kModeRelabelNew) | mode_bit(kModeRelabelCont);
constexpr unsigned kGateDischarge = mode_bit(kModeDischarge);
constexpr unsigned kGateClosureNew = mode_bit(kModeClosureNew) | mode_bit(kModeFinalNew);
constexpr unsigned kGateClosure =
    mode_bit(kModeClosureNew) | mode_bit(kModeClosureCont) | mode_bit(kModeFinalNew) | mode_bit(kModeFinalCont);
constexpr unsigned kGateInit = mode_bit(kModeInit);
// with this bit, a closure-stage gate also closes once the termination test has reached t
constexpr unsigned kGateUnlessReached = 1u << 31;

struct SolveControl {
	int32_t mode;
	int32_t rounds;      // global relabels completed
	int32_t reached;     // the termination closure put a voxel with a residual sink link into S
	int32_t status;      // init status (bit 0 negative capacity, bit 1 int32 headroom, bit 2 int16 staging)
	int32_t first[2];    // tile counts of the current first sweep's two colours (0 when not starting one)
	int32_t dirty;       // dirty tiles at the last stage check
	int32_t active;      // active tiles at the last relabel check
	int32_t phases;      // discharge phases run
	int32_t max_rounds;  // round cap (kModeNotConverged beyond it)
	int32_t pad[6];
};
static_assert(sizeof(SolveControl) == 64, "the control block is read back as 64 bytes");

__device__ inline bool closure_check_mode(int m) { return m == kModeClosureNew || m == kModeClosureCont; }

// A kernel taking part in a device-driven solve runs only while the control block's stage is in
// its gate mask; ctl null (host-driven solve) means always.
__device__ inline bool gate_open(const SolveControl *ctl, unsigned mask)
{
	if (!ctl) {
		return true;
	}
	const int m = *(volatile const int32_t *)&ctl->mode;
	if (!((mask >> m) & 1u)) {
		return false;
	}
	if ((mask & kGateUnlessReached) && closure_check_mode(m) && *(volatile const int32_t *)&ctl->reached) {
		return false;
	}
	return true;
}

}  // namespace cgp
