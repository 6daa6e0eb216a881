// Learned cut proposal: a small 3D CNN (4 conv layers 3x3x3, 16 channels, ReLU, sigmoid head)
// run with BF16 operands and FP32 accumulation on the matrix cores. Its thresholded output is a
// candidate source side; the capacity of that cut bounds the max flow from above, so a solve may
// stop as soon as the flow into t reaches it (max-flow = min-cut), and runs to completion otherwise.
#pragma once

#include <cstdint>
#include <vector>

#include "common.h"

namespace cgp {

class GridSolver;

// Parameter vector (fp32), in this order:
//   W1[27][4][16]  (tap, input channel, output channel; input channel 3 is padding, kept at 0)
//   b1[16]
//   W2, W3, W4 [27][16][16] with b2, b3, b4 [16] after each
//   w_head[16], b_head
// Taps enumerate (dz, dy, dx) in {-1, 0, 1}^3 with dx fastest: t = (dz+1)*9 + (dy+1)*3 + (dx+1).
constexpr int kCnnChannels = 16;
constexpr int kCnnTaps = 27;
constexpr int kCnnW1 = kCnnTaps * 4 * kCnnChannels;
constexpr int kCnnW = kCnnTaps * kCnnChannels * kCnnChannels;
constexpr int kCnnParams = kCnnW1 + kCnnChannels + 3 * (kCnnW + kCnnChannels) + kCnnChannels + 1;

class CnnModel {
public:
	explicit CnnModel(hipStream_t stream);
	void set_params(const float *params);
	void get_params(float *params);
	// Forward on grid 0 of `grid` with the given volume (host or device bytes). Keeps the logits and,
	// if `keep_activations`, every layer's activations for a backward pass.
	void forward(GridSolver &grid, const uint8_t *volume, bool device_volume, bool keep_activations);
	// Thresholded prediction (logit > 0) as a 0/1 byte mask in device memory of the grid's size.
	const uint8_t *predicted_side();
	void prediction_to_host(uint8_t *side, float *prob);
	// Gradient of the mean binary cross-entropy against the grid's S_min (the grid must be solved);
	// returns the loss. grad (host, kCnnParams) may be null.
	double gradients(GridSolver &grid, const uint8_t *volume, bool device_volume, float *grad);
	// One Adam step on that gradient; returns the loss before the step.
	double train_step(GridSolver &grid, const uint8_t *volume, bool device_volume, float learning_rate);
	int64_t voxels() const { return mVox; }

private:
	void upload_packed();
	void ensure(int64_t vox);
	void run_backward(GridSolver &grid, double &loss);

	hipStream_t mStream;
	int mNx = 0, mNy = 0, mNz = 0;
	int64_t mVox = 0;
	std::vector<float> mHost;                     // fp32 master parameters (host mirror)
	DeviceBuffer<float> mParams, mGrad, mAdamM, mAdamV, mPartials;
	DeviceBuffer<uint16_t> mPackFwd, mPackBwd;    // bf16 weight fragments, MFMA operand order
	DeviceBuffer<uint16_t> mX0;                   // input, 4 channels
	DeviceBuffer<uint16_t> mAct[4];               // post-ReLU activations of the 4 conv layers
	DeviceBuffer<uint16_t> mG[2];                 // pre-activation gradients (ping-pong)
	DeviceBuffer<float> mLogit;
	DeviceBuffer<uint8_t> mVolume, mSide;
	DeviceBuffer<double> mScalars;                // loss and head-bias gradient accumulators
	int mStep = 0;
	bool mPackDirty = true;
};

}  // namespace cgp
