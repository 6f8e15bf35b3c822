// ipf_learned.h -- sizes and launchers of the learned return-branch
// predictor (ipf_learned.hip).
#pragma once

#include <cstdint>

#include "ipf_kernels.h"

namespace ipf {

namespace mlp {

constexpr int IN = 7;            // features
constexpr int INP = 16;          // padded to one MFMA K step
constexpr int H = 64;            // hidden width
constexpr int NBR = 5;           // branch logits
constexpr int OUT = NBR + 1;     // + plastic increment
constexpr int OUTP = 16;         // padded to one MFMA tile
constexpr int FEAT_STRIDE = 8;   // floats per sample in a feature array
constexpr float REG_WEIGHT = 0.1f;

// FP32 parameter vector (padded entries stay zero): W1 [INP x H], b1 [H],
// W2 [H x H], b2 [H], W3 [H x OUTP], b3 [OUTP]; row-major (input, output).
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + INP * H;
constexpr int OFF_W2 = OFF_B1 + H;
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H * OUTP;
constexpr int NPAR = OFF_B3 + OUTP;
constexpr int NPAR_TRAINABLE = IN * H + H + H * H + H + H * OUT + OUT;  // 5062

// gradient of the mean loss over n samples into per-wavefront slots
// (nblocks * 4 slots of NPAR floats) and summed losses (nblocks * 4 doubles)
void launch_grad(const short* wbf, const float* par, const float* feat, const int* label, const float* target,
                 int64_t n, int nblocks, float* slots, double* loss_slots, hipStream_t s);
// grad = sum of the slots in order; with adam, one Adam step on par / m1 / m2
// (bias corrections bc1 = 1 - b1^t, bc2 = 1 - b2^t) and BF16 copies in wbf
void launch_reduce(const float* slots, int nslots, float* grad, float* par, float* m1, float* m2, short* wbf,
                   bool adam, float lr, float b1, float b2, float eps, float bc1, float bc2, hipStream_t s);
void launch_refresh(const float* par, short* wbf, hipStream_t s);
void launch_predict(const short* wbf, const float* par, const float* feat, int64_t n, int nblocks, int* branch,
                    float* out, hipStream_t s);

}  // namespace mlp

// Features [npts x FEAT_STRIDE], exact branch and 1e3 * d_alpha of every
// point of one group at the trial state (nothing committed); reference
// point order through orig.
void launch_point_features(int shape, const GroupDev& g, const double* strain, const int64_t* orig,
                           const PointState& st, float* feat, int* label, float* target, hipStream_t stream);

// Stable counting sort of n keys in [0, IPF_NBRANCH) (out-of-range keys
// sort last): order[] = indexes grouped by key, index order inside a key.
// hist holds key_sort_blocks(n) * IPF_NBRANCH ints.
constexpr int IPF_NBRANCH = 5;
constexpr int IPF_SORT_BLOCK = 256;  // keys per histogram block
int key_sort_blocks(int64_t n);
void launch_branch_order(const int* key, int64_t n, int* hist, int* order, hipStream_t stream);
// The exact return map of the points in `order` (reference indexes; new_of
// maps them to state slots, pmat to material ids).
void launch_sigma_order(int64_t n, const int* order, const int64_t* new_of, const int* pmat, const MatConst* mats,
                        const double* strain, double* wsigma, const PointState& st, bool commit, hipStream_t stream);

}  // namespace ipf
