// Learned carrier-mirror mask: network state on one device, forward,
// gradient, Adam and the weights file (chd_mask_net_* in
// include/chromadec/tpal_mask.h).  The kernels are in mask_net_kernels.hip.
#include <chromadec/tpal_mask.h>

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "mask_net_internal.h"
#include "mask_net_state.h"

namespace {

std::string hip_msg(const char *what, hipError_t e)
{
    return std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
}

constexpr char kMagic[8] = {'C', 'H', 'D', 'M', 'A', 'S', 'K', '1'};

void release(chd_mask_net *n)
{
    void *ptrs[] = {n->theta, n->adam_m, n->adam_v, n->partial, n->mask, n->f.w1, n->f.w2, n->f.w2t, n->f.w3t,
                    n->f.w3, n->a.feat, n->a.act1, n->a.act2, n->a.dpre1, n->a.dpre2, n->a.logit, n->a.w,
                    n->a.num, n->a.dlogit, n->a.loss_tile};
    for (void *p : ptrs)
        if (p) hipFree(p);
}

}  // namespace

extern "C" {

chd_status_t chd_mask_net_create(int32_t device, int64_t max_tiles, chd_mask_net_t **out)
{
    static const char *W = "chd_mask_net_create";
    if (out == nullptr || max_tiles <= 0 || max_tiles > (1 << 22))
        return chd::fail(CHD_E_INVALID_ARG, std::string(W) + ": null output or max_tiles outside 1..4194304");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return chd::fail(CHD_E_UNSUPPORTED, std::string(W) + ": no HIP device");
    if (device < 0 || device >= ndev) return chd::fail(CHD_E_INVALID_ARG, std::string(W) + ": bad device ordinal");
    hipSetDevice(device);
    auto *n = new chd_mask_net();
    n->device = device;
    n->max_tiles = max_tiles;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        n->cus = cus;
    const size_t T = (size_t)max_tiles, P = mnet::kP;
    bool ok = true;
    auto dm = [&](void **p, size_t bytes) {
        if (!ok) return;
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            ok = false;
        }
    };
    dm((void **)&n->theta, mnet::kNParams * sizeof(float));
    dm((void **)&n->adam_m, mnet::kNParams * sizeof(float));
    dm((void **)&n->adam_v, mnet::kNParams * sizeof(float));
    dm((void **)&n->partial, (size_t)mnet::kWgradSlots * mnet::kNParams * sizeof(float));
    dm((void **)&n->mask, T * P * sizeof(float));
    dm((void **)&n->f.w1, mnet::kK1 * 64 * sizeof(uint4));
    dm((void **)&n->f.w2, mnet::kK2 * 64 * sizeof(uint4));
    dm((void **)&n->f.w2t, mnet::kK2 * 64 * sizeof(uint4));
    dm((void **)&n->f.w3t, mnet::kK3 * 64 * sizeof(uint4));
    dm((void **)&n->f.w3, mnet::kK3 * 64 * sizeof(uint4));
    dm((void **)&n->a.feat, T * P * 4 * sizeof(uint16_t));
    dm((void **)&n->a.act1, T * P * mnet::kCh * sizeof(uint16_t));
    dm((void **)&n->a.act2, T * P * mnet::kCh * sizeof(uint16_t));
    dm((void **)&n->a.dpre1, T * P * mnet::kCh * sizeof(uint16_t));
    dm((void **)&n->a.dpre2, T * P * mnet::kCh * sizeof(uint16_t));
    dm((void **)&n->a.logit, T * P * sizeof(float));
    dm((void **)&n->a.w, T * P * sizeof(float));
    dm((void **)&n->a.num, T * P * sizeof(float));
    dm((void **)&n->a.dlogit, T * P * sizeof(float));
    dm((void **)&n->a.loss_tile, T * sizeof(float));
    if (ok) {
        ok = hipMemset(n->theta, 0, mnet::kNParams * sizeof(float)) == hipSuccess &&
             hipMemset(n->adam_m, 0, mnet::kNParams * sizeof(float)) == hipSuccess &&
             hipMemset(n->adam_v, 0, mnet::kNParams * sizeof(float)) == hipSuccess;
    }
    if (ok) {
        mnet::launch_arrange(n->theta, n->f, nullptr);
        ok = hipDeviceSynchronize() == hipSuccess;
    }
    if (!ok) {
        release(n);
        delete n;
        return chd::fail(CHD_E_INTERNAL, std::string(W) + ": device allocation failed");
    }
    *out = n;
    return CHD_OK;
}

void chd_mask_net_free(chd_mask_net_t *net)
{
    if (net == nullptr) return;
    hipSetDevice(net->device);
    hipDeviceSynchronize();
    release(net);
    delete net;
}

chd_status_t chd_mask_net_set_params(chd_mask_net_t *net, const float *params, int64_t n)
{
    if (net == nullptr || params == nullptr || n != mnet::kNParams)
        return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_set_params: null argument or wrong parameter count");
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i]))
            return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_set_params: non-finite parameter");
    hipSetDevice(net->device);
    hipError_t e = hipMemcpy(net->theta, params, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        mnet::launch_arrange(net->theta, net->f, nullptr);
        e = hipDeviceSynchronize();
    }
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_mask_net_set_params", e));
    return CHD_OK;
}

chd_status_t chd_mask_net_get_params(chd_mask_net_t *net, float *params, int64_t n)
{
    if (net == nullptr || params == nullptr || n != mnet::kNParams)
        return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_get_params: null argument or wrong parameter count");
    hipSetDevice(net->device);
    const hipError_t e = hipMemcpy(params, net->theta, n * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_mask_net_get_params", e));
    return CHD_OK;
}

chd_status_t chd_mask_net_load(chd_mask_net_t *net, const char *path)
{
    if (net == nullptr || path == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_load: null argument");
    FILE *fp = std::fopen(path, "rb");
    if (fp == nullptr) return chd::fail(CHD_E_FILE_NOT_FOUND, std::string("chd_mask_net_load: cannot open ") + path);
    char magic[8];
    int32_t count = 0;
    std::vector<float> p(mnet::kNParams);
    const bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
                    std::fread(&count, 4, 1, fp) == 1 && count == mnet::kNParams &&
                    std::fread(p.data(), sizeof(float), p.size(), fp) == p.size();
    std::fclose(fp);
    if (!ok) return chd::fail(CHD_E_NN_MODEL_LOAD, std::string("chd_mask_net_load: not a mask weights file: ") + path);
    return chd_mask_net_set_params(net, p.data(), mnet::kNParams);
}

chd_status_t chd_mask_net_save(chd_mask_net_t *net, const char *path)
{
    if (net == nullptr || path == nullptr) return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_save: null argument");
    std::vector<float> p(mnet::kNParams);
    chd_status_t st = chd_mask_net_get_params(net, p.data(), mnet::kNParams);
    if (st != CHD_OK) return st;
    FILE *fp = std::fopen(path, "wb");
    if (fp == nullptr) return chd::fail(CHD_E_IO, std::string("chd_mask_net_save: cannot write ") + path);
    const int32_t count = mnet::kNParams;
    const bool ok = std::fwrite(kMagic, 1, 8, fp) == 8 && std::fwrite(&count, 4, 1, fp) == 1 &&
                    std::fwrite(p.data(), sizeof(float), p.size(), fp) == p.size();
    const bool closed = std::fclose(fp) == 0;
    if (!ok || !closed) return chd::fail(CHD_E_IO, std::string("chd_mask_net_save: write failed: ") + path);
    return CHD_OK;
}

chd_status_t chd_mask_net_forward(chd_mask_net_t *net, const void *d_spectra, int64_t n_tiles, float *d_mask,
                                  void *stream)
{
    if (net == nullptr || n_tiles < 0 || (n_tiles > 0 && (d_spectra == nullptr || d_mask == nullptr)))
        return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_forward: null argument");
    hipSetDevice(net->device);
    const hipStream_t s = (hipStream_t)stream;
    // fused kernel: no working memory, any number of tiles in one launch
    mnet::launch_infer(reinterpret_cast<const float2 *>(d_spectra), n_tiles, net->theta, net->f, d_mask, net->cus, s);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_mask_net_forward", e));
    return CHD_OK;
}

chd_status_t chd_mask_net_gradient(chd_mask_net_t *net, const void *d_spectra, const void *d_truth_spectra,
                                   int64_t n_tiles, float loss_scale, float *d_grad, double *loss, void *stream)
{
    if (net == nullptr || n_tiles <= 0 || d_spectra == nullptr || d_truth_spectra == nullptr || d_grad == nullptr)
        return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_gradient: null argument or no tiles");
    if (n_tiles > net->max_tiles)
        return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_gradient: more tiles than the network was created for");
    if (!std::isfinite(loss_scale)) return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_gradient: bad loss scale");
    hipSetDevice(net->device);
    const hipStream_t s = (hipStream_t)stream;
    const float2 *spec = reinterpret_cast<const float2 *>(d_spectra);
    mnet::launch_features(spec, n_tiles, net->a.feat, s);
    mnet::launch_targets(spec, reinterpret_cast<const float2 *>(d_truth_spectra), n_tiles, net->a.w, net->a.num, s);
    mnet::launch_forward(net->theta, net->f, net->a, n_tiles, net->mask, s);
    mnet::launch_backward(net->theta, net->f, net->a, net->mask, n_tiles, loss_scale, net->partial, d_grad, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && loss != nullptr) {
        std::vector<float> lt((size_t)n_tiles);
        e = hipMemcpyAsync(lt.data(), net->a.loss_tile, n_tiles * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        double acc = 0.0;
        for (float v : lt) acc += v;
        *loss = acc;
    }
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_mask_net_gradient", e));
    return CHD_OK;
}

chd_status_t chd_mask_net_adam_step(chd_mask_net_t *net, const float *d_grad, float lr, void *stream)
{
    if (net == nullptr || d_grad == nullptr || !(lr > 0.0f))
        return chd::fail(CHD_E_INVALID_ARG, "chd_mask_net_adam_step: null argument or learning rate <= 0");
    hipSetDevice(net->device);
    const hipStream_t s = (hipStream_t)stream;
    net->step += 1;
    mnet::launch_adam(net->theta, d_grad, net->adam_m, net->adam_v, lr, 0.9f, 0.999f, 1e-8f, net->step, s);
    mnet::launch_arrange(net->theta, net->f, s);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return chd::fail(CHD_E_INTERNAL, hip_msg("chd_mask_net_adam_step", e));
    return CHD_OK;
}

}  // extern "C"
