// ipf_mlp.hip -- host side of the learned return-branch predictor: the
// ipf_mlp object of include/intpointsfem.h (parameters, Adam moments, BF16
// weight copies and gradient slots in device memory).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/intpointsfem.h"
#include "ipf_learned.h"

using namespace ipf;

extern "C" int ipf_internal_set_error(const char* msg);

struct ipf_mlp {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;  // own_stream or the caller's
    float* par = nullptr;     // FP32 master parameters [NPAR]
    float* m1 = nullptr;      // Adam moments
    float* m2 = nullptr;
    short* wbf = nullptr;     // BF16 copies
    float* grad = nullptr;    // last gradient [NPAR]
    float* slots = nullptr;   // [kMaxBlocks * 4][NPAR]
    double* loss_slots = nullptr;
    int64_t step = 0;
};

namespace {

constexpr int kMaxBlocks = 256;

int mlp_fail(int code, const char* msg) {
    ipf_internal_set_error(msg);
    return code;
}

int hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return IPF_OK;
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return mlp_fail(IPF_EHIP, m.c_str());
}

#define MLP_HIP(call)                                   \
    do {                                                \
        int rc_ = hip_check((call), #call);             \
        if (rc_) return rc_;                            \
    } while (0)

int grad_blocks(int64_t n) {
    const int64_t tiles = (n + 15) / 16;
    const int64_t b = (tiles + 3) / 4;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, kMaxBlocks));
}

// splitmix64: the initial weights depend only on the seed
uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void init_params(std::vector<float>& p, uint64_t seed) {
    using namespace mlp;
    p.assign(NPAR, 0.f);
    uint64_t st = seed;
    auto uni = [&]() { return (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0); };
    // Glorot-uniform weights, zero biases; padded rows / columns stay zero
    auto fill = [&](int off, int rows, int cols, int ld, int fan_in, int fan_out) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < rows; i++)
            for (int j = 0; j < cols; j++) p[off + i * ld + j] = (float)((2.0 * uni() - 1.0) * a);
    };
    fill(OFF_W1, IN, H, H, IN, H);
    fill(OFF_W2, H, H, H, H, H);
    fill(OFF_W3, H, OUT, OUTP, H, OUT);
}

}  // namespace

extern "C" {

int ipf_mlp_create(int device, uint64_t seed, ipf_mlp** out) {
    using namespace mlp;
    if (!out) return mlp_fail(IPF_EINVAL, "ipf_mlp_create: NULL out");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return mlp_fail(IPF_EHIP, "ipf_mlp_create: no HIP device available");
    if (device < 0 || device >= ndev) return mlp_fail(IPF_EINVAL, "ipf_mlp_create: bad device");
    MLP_HIP(hipSetDevice(device));
    auto* m = new ipf_mlp();
    m->device = device;
    auto bad = [&](int rc) {
        ipf_mlp_destroy(m);
        return rc;
    };
    int rc;
    if ((rc = hip_check(hipStreamCreate(&m->own_stream), "hipStreamCreate")) ||
        (rc = hip_check(hipMalloc((void**)&m->par, sizeof(float) * NPAR), "hipMalloc")) ||
        (rc = hip_check(hipMalloc((void**)&m->m1, sizeof(float) * NPAR), "hipMalloc")) ||
        (rc = hip_check(hipMalloc((void**)&m->m2, sizeof(float) * NPAR), "hipMalloc")) ||
        (rc = hip_check(hipMalloc((void**)&m->wbf, sizeof(short) * NPAR), "hipMalloc")) ||
        (rc = hip_check(hipMalloc((void**)&m->grad, sizeof(float) * NPAR), "hipMalloc")) ||
        (rc = hip_check(hipMalloc((void**)&m->slots, sizeof(float) * NPAR * 4 * kMaxBlocks), "hipMalloc")) ||
        (rc = hip_check(hipMalloc((void**)&m->loss_slots, sizeof(double) * 4 * kMaxBlocks), "hipMalloc")))
        return bad(rc);
    m->stream = m->own_stream;
    std::vector<float> p;
    init_params(p, seed);
    if ((rc = ipf_mlp_set_params(m, p.data()))) return bad(rc);
    *out = m;
    return IPF_OK;
}

int ipf_mlp_destroy(ipf_mlp* m) {
    if (!m) return IPF_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : {(void*)m->par, (void*)m->m1, (void*)m->m2, (void*)m->wbf, (void*)m->grad, (void*)m->slots,
                    (void*)m->loss_slots})
        if (p) (void)hipFree(p);
    if (m->own_stream) (void)hipStreamDestroy(m->own_stream);
    delete m;
    return IPF_OK;
}

int ipf_mlp_set_stream(ipf_mlp* m, void* stream) {
    if (!m) return mlp_fail(IPF_EINVAL, "ipf_mlp_set_stream: NULL mlp");
    m->stream = stream ? (hipStream_t)stream : m->own_stream;
    return IPF_OK;
}

int ipf_mlp_num_params(int* padded, int* trainable) {
    if (padded) *padded = mlp::NPAR;
    if (trainable) *trainable = mlp::NPAR_TRAINABLE;
    return IPF_OK;
}

int ipf_mlp_set_params(ipf_mlp* m, const float* params_host) {
    using namespace mlp;
    if (!m || !params_host) return mlp_fail(IPF_EINVAL, "ipf_mlp_set_params: NULL argument");
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(hipMemcpyAsync(m->par, params_host, sizeof(float) * NPAR, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(hipMemsetAsync(m->m1, 0, sizeof(float) * NPAR, m->stream));
    MLP_HIP(hipMemsetAsync(m->m2, 0, sizeof(float) * NPAR, m->stream));
    launch_refresh(m->par, m->wbf, m->stream);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(m->stream));
    m->step = 0;
    return IPF_OK;
}

int ipf_mlp_get_params(ipf_mlp* m, float* params_host) {
    if (!m || !params_host) return mlp_fail(IPF_EINVAL, "ipf_mlp_get_params: NULL argument");
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(hipStreamSynchronize(m->stream));
    MLP_HIP(hipMemcpy(params_host, m->par, sizeof(float) * mlp::NPAR, hipMemcpyDeviceToHost));
    return IPF_OK;
}

static int run_grad(ipf_mlp* m, const float* feat, const int* branch, const float* target, int64_t n, bool adam,
                    float lr, double* loss) {
    using namespace mlp;
    if (!m || !feat || !branch || !target) return mlp_fail(IPF_EINVAL, "ipf_mlp: NULL argument");
    if (n < 1) return mlp_fail(IPF_EINVAL, "ipf_mlp: need n >= 1 samples");
    MLP_HIP(hipSetDevice(m->device));
    const int nb = grad_blocks(n);
    launch_grad(m->wbf, m->par, feat, branch, target, n, nb, m->slots, m->loss_slots, m->stream);
    float bc1 = 1.f, bc2 = 1.f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam) {
        m->step++;
        bc1 = 1.f - std::pow(b1, (float)m->step);
        bc2 = 1.f - std::pow(b2, (float)m->step);
    }
    launch_reduce(m->slots, nb * 4, m->grad, m->par, m->m1, m->m2, m->wbf, adam, lr, b1, b2, 1e-8f, bc1, bc2,
                  m->stream);
    MLP_HIP(hipGetLastError());
    if (loss) {
        std::vector<double> ls(nb * 4);
        MLP_HIP(hipMemcpyAsync(ls.data(), m->loss_slots, sizeof(double) * nb * 4, hipMemcpyDeviceToHost, m->stream));
        MLP_HIP(hipStreamSynchronize(m->stream));
        double s = 0.0;
        for (double v : ls) s += v;
        *loss = s / (double)n;
    }
    return IPF_OK;
}

int ipf_mlp_gradient(ipf_mlp* m, const float* feat_dev, const int* branch_dev, const float* target_dev, int64_t n,
                     float* grad_host, double* loss) {
    int rc = run_grad(m, feat_dev, branch_dev, target_dev, n, false, 0.f, loss);
    if (rc) return rc;
    if (grad_host) {
        MLP_HIP(hipMemcpyAsync(grad_host, m->grad, sizeof(float) * mlp::NPAR, hipMemcpyDeviceToHost, m->stream));
        MLP_HIP(hipStreamSynchronize(m->stream));
    }
    return IPF_OK;
}

int ipf_mlp_train_step(ipf_mlp* m, const float* feat_dev, const int* branch_dev, const float* target_dev, int64_t n,
                       float lr, double* loss) {
    if (!(lr > 0.f)) return mlp_fail(IPF_EINVAL, "ipf_mlp_train_step: need lr > 0");
    return run_grad(m, feat_dev, branch_dev, target_dev, n, true, lr, loss);
}

int ipf_mlp_predict(ipf_mlp* m, const float* feat_dev, int64_t n, int* branch_dev, float* out_dev) {
    if (!m || !feat_dev || (!branch_dev && !out_dev)) return mlp_fail(IPF_EINVAL, "ipf_mlp_predict: NULL argument");
    if (n < 0) return mlp_fail(IPF_EINVAL, "ipf_mlp_predict: n < 0");
    if (n == 0) return IPF_OK;
    MLP_HIP(hipSetDevice(m->device));
    const int64_t tiles = (n + 15) / 16;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 2048));
    mlp::launch_predict(m->wbf, m->par, feat_dev, n, nb, branch_dev, out_dev, m->stream);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(m->stream));
    return IPF_OK;
}

}  // extern "C"
