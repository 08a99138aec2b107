// C ABI of the learned-closure trainer (tc_mlp_*): FP32 master weights, Adam
// state and bfloat16 device copies on one GPU. No reference counterpart.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "kernels.h"
#include "mlp.h"

using namespace tc;

void tc_set_error(const char* msg);  // engine.cpp

struct tc_mlp {
    int dev = 0;
    hipStream_t s = nullptr;
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    unsigned short* w = nullptr;
    float* b = nullptr;
    float* partial = nullptr;
    size_t partial_recs = 0;
    double* tmp = nullptr;
    double* loss = nullptr;
    float* scratch = nullptr;  // this device's copy of the no-allocation scratch (set at create)
    long long steps = 0;
    MlpDev dev_view() const {
        MlpDev d;
        d.w = w;
        d.b = b;
        return d;
    }
};

// Scratch of tc_mlp_train_step_noalloc: the gradient kernel's per-block partial
// records, as many as that kernel ever launches blocks for (kMlpTrainMaxBlocks x
// kMlpRec floats = 19.3 MB), so every batch size fits. Each GPU that loads this
// library holds one copy until the process ends; host code reaches it only through
// the address hipGetSymbolAddress returns on that device.
constexpr size_t kNoAllocRecs = kMlpTrainMaxBlocks;
static __device__ __align__(256) float scratch[kNoAllocRecs * kMlpRec];

namespace {

constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The scratch's address on `device`, looked up once per device; the device must be current.
hipError_t scratch_address(int device, float** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = (float*)p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

#define MCHK(x)                                                                      \
    do {                                                                             \
        hipError_t _e = (x);                                                         \
        if (_e != hipSuccess) {                                                      \
            char _b[256];                                                            \
            snprintf(_b, sizeof _b, "HIP error %s at %s:%d", hipGetErrorString(_e), \
                     __FILE__, __LINE__);                                            \
            tc_set_error(_b);                                                        \
            return TC_ERR_HIP;                                                       \
        }                                                                            \
    } while (0)

int arg_error(const char* m) {
    tc_set_error(m);
    return TC_ERR_ARG;
}

uint64_t splitmix(uint64_t& st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases, seeded.
void init_params(unsigned long long seed, std::vector<float>& p) {
    p.assign(kMlpParams, 0.f);
    uint64_t st = seed;
    auto fill = [&](int off, int n, int fan_in) {
        const double b = 1.0 / std::sqrt((double)fan_in);
        for (int i = 0; i < n; i++) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            p[off + i] = (float)((2.0 * u - 1.0) * b);
        }
    };
    fill(kOffW1, kMlpHid * kMlpIn, kMlpIn);
    fill(kOffB1, kMlpHid, kMlpIn);
    fill(kOffW2, kMlpHid * kMlpHid, kMlpHid);
    fill(kOffB2, kMlpHid, kMlpHid);
    fill(kOffW3, kMlpHid * kMlpHid, kMlpHid);
    fill(kOffB3, kMlpHid, kMlpHid);
    fill(kOffW4, kMlpOut * kMlpHid, kMlpHid);
    fill(kOffB4, kMlpOut, kMlpHid);
}

int upload(tc_mlp* m, const float* p) {
    std::vector<unsigned short> w;
    std::vector<float> b;
    mlp_pack_device(p, w, b);
    MCHK(hipMemcpy(m->params, p, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(m->w, w.data(), w.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(m->b, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemset(m->m, 0, kMlpParams * sizeof(float)));
    MCHK(hipMemset(m->v, 0, kMlpParams * sizeof(float)));
    m->steps = 0;
    return TC_OK;
}

int ensure_partial(tc_mlp* m, long long n) {
    const size_t recs = (size_t)mlp_train_blocks(n);
    if (recs <= m->partial_recs) return TC_OK;
    if (m->partial) MCHK(hipFree(m->partial));
    m->partial = nullptr;
    MCHK(hipMalloc(&m->partial, recs * kMlpRec * sizeof(float)));
    m->partial_recs = recs;
    return TC_OK;
}

// Gradient (and Adam step) on n samples. The partial records go to the trainer's
// own buffer (grown on demand), or with no_alloc to this device's static scratch.
int grad_and_maybe_adam(tc_mlp* m, const float* x, const float* t, long long n, const AdamArgs& a, double* loss,
                        bool no_alloc = false) {
    if (!m) return arg_error("null trainer");
    if (n <= 0 || !x || !t) return arg_error("need n > 0 samples and device inputs/targets");
    MCHK(hipSetDevice(m->dev));
    float* partial = nullptr;
    if (no_alloc) {
        if (n > (long long)INT_MAX) return arg_error("no-allocation training step: n must be at most 2^31 - 1 samples");
        if ((size_t)mlp_train_blocks(n) > kNoAllocRecs)
            return arg_error("no-allocation training step: batch needs more partial records than the scratch holds");
        if (!m->scratch) return arg_error("no-allocation training step: no scratch on this device");
        partial = m->scratch;
    } else {
        if (int r = ensure_partial(m, n)) return r;
        partial = m->partial;
    }
    MCHK(launch_mlp_partials(m->s, m->dev_view(), x, t, n, partial));
    MCHK(launch_mlp_reduce(m->s, partial, mlp_train_blocks(n), m->tmp, m->grad, m->loss, a, m->params, m->m, m->v,
                           m->w, m->b));
    double l = 0;
    MCHK(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->s));
    MCHK(hipStreamSynchronize(m->s));
    if (loss) *loss = l / ((double)n * kMlpOut);
    if (!std::isfinite(l)) {
        tc_set_error("non-finite training loss");
        return TC_ERR_NUMERIC;
    }
    return TC_OK;
}

}  // namespace

extern "C" {

int tc_mlp_param_count(void) { return kMlpParams; }

int tc_mlp_create(int device, unsigned long long seed, tc_mlp** out) {
    if (!out) return arg_error("null output handle");
    *out = nullptr;
    int ndev = 0;
    MCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return arg_error("no such device");
    MCHK(hipSetDevice(device));
    tc_mlp* m = new tc_mlp();
    m->dev = device;
    auto fail = [&](int r) {
        tc_mlp_destroy(m);
        return r;
    };
    if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess) return fail(TC_ERR_HIP);
    if (device < kMaxDevices) {  // the no-allocation step's scratch on this device
        const hipError_t e = scratch_address(device, &m->scratch);
        if (e != hipSuccess) {
            tc_set_error("hipGetSymbolAddress of the training scratch failed");
            return fail(TC_ERR_HIP);
        }
    }
    if (hipMalloc(&m->params, kMlpParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->m, kMlpParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->v, kMlpParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->grad, kMlpParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->w, kWBlock * sizeof(unsigned short)) != hipSuccess ||
        hipMalloc(&m->b, kBBlock * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->tmp, (size_t)kMlpReduceSplit * kMlpRec * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->loss, sizeof(double)) != hipSuccess) {
        tc_set_error("device allocation failed");
        return fail(TC_ERR_HIP);
    }
    std::vector<float> p;
    init_params(seed, p);
    if (int r = upload(m, p.data())) return fail(r);
    *out = m;
    return TC_OK;
}

int tc_mlp_destroy(tc_mlp* m) {
    if (!m) return TC_OK;
    hipSetDevice(m->dev);
    if (m->s) hipStreamSynchronize(m->s);
    for (void* p : {(void*)m->params, (void*)m->m, (void*)m->v, (void*)m->grad, (void*)m->w, (void*)m->b,
                    (void*)m->partial, (void*)m->tmp, (void*)m->loss})
        if (p) hipFree(p);
    if (m->s) hipStreamDestroy(m->s);
    delete m;
    return TC_OK;
}

int tc_mlp_set_params(tc_mlp* m, const float* host, size_t count) {
    if (!m || !host) return arg_error("null argument");
    if (count != (size_t)kMlpParams) return arg_error("parameter count mismatch");
    MCHK(hipSetDevice(m->dev));
    MCHK(hipStreamSynchronize(m->s));
    return upload(m, host);
}

int tc_mlp_get_params(tc_mlp* m, float* host, size_t count) {
    if (!m || !host) return arg_error("null argument");
    if (count != (size_t)kMlpParams) return arg_error("parameter count mismatch");
    MCHK(hipSetDevice(m->dev));
    MCHK(hipMemcpyAsync(host, m->params, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, m->s));
    MCHK(hipStreamSynchronize(m->s));
    return TC_OK;
}

int tc_mlp_forward(tc_mlp* m, const float* dev_x, long long n, float* dev_y) {
    if (!m || (n > 0 && (!dev_x || !dev_y)) || n < 0) return arg_error("bad forward arguments");
    MCHK(hipSetDevice(m->dev));
    MCHK(launch_mlp_forward(m->s, m->dev_view(), dev_x, n, dev_y));
    MCHK(hipStreamSynchronize(m->s));
    return TC_OK;
}

int tc_mlp_gradient(tc_mlp* m, const float* dev_x, const float* dev_t, long long n, float* host_grad, size_t count,
                    double* loss) {
    if (host_grad && count != (size_t)kMlpParams) return arg_error("parameter count mismatch");
    AdamArgs a;
    a.apply = 0;
    if (int r = grad_and_maybe_adam(m, dev_x, dev_t, n, a, loss)) return r;
    if (host_grad) {
        MCHK(hipMemcpyAsync(host_grad, m->grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, m->s));
        MCHK(hipStreamSynchronize(m->s));
    }
    return TC_OK;
}

static int train_step(tc_mlp* m, const float* dev_x, const float* dev_t, long long n, float lr, double* loss,
                      bool no_alloc) {
    if (!m) return arg_error("null trainer");
    if (!(lr > 0.f)) return arg_error("learning rate must be positive");
    AdamArgs a;
    a.apply = 1;
    a.lr = lr;
    const long long step = m->steps + 1;
    a.bc1 = (float)(1.0 - std::pow((double)a.b1, (double)step));
    a.bc2 = (float)(1.0 - std::pow((double)a.b2, (double)step));
    if (int r = grad_and_maybe_adam(m, dev_x, dev_t, n, a, loss, no_alloc)) return r;
    m->steps = step;
    return TC_OK;
}

int tc_mlp_train_step(tc_mlp* m, const float* dev_x, const float* dev_t, long long n, float lr, double* loss) {
    return train_step(m, dev_x, dev_t, n, lr, loss, false);
}

int tc_mlp_train_step_noalloc(tc_mlp* m, const float* dev_x, const float* dev_t, long long n, float lr,
                              double* loss) {
    return train_step(m, dev_x, dev_t, n, lr, loss, true);
}

int tc_mlp_steps(const tc_mlp* m, long long* steps) {
    if (!m || !steps) return arg_error("null argument");
    *steps = m->steps;
    return TC_OK;
}

}  // extern "C"
