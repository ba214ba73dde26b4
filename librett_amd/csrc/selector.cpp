// Learned plan selector: features, loaded weights, GPU inference for
// librettPlan, and the C-ABI of the MLP kernels (training and inference).
//
// The reference chooses among candidates with its analytic model only
// (choosePlanHeuristic, src/plan.cpp:721-789); its librettPlanMeasure
// (src/librett.cpp:224-343) times every candidate, which is where the
// selector's training labels come from.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "devmem.h"
#include "../../include/librett_c.h"
#include "selector.h"
#include "selector_ops.h"

namespace lrt {

static_assert(mlp::F == kSelFeatures, "feature count");

namespace {

float lg(double v) { return v > 0 ? (float)std::log2(v) : 0.0f; }

struct DevBuf {
  float* params = nullptr;
  float* x = nullptr;
  float* out = nullptr;
  int cap = 0;
  uint64_t gen = 0;
  hipStream_t stream = nullptr;
  bool warm = false;  // the inference kernel has run once on this device
};

// Weights and per-device inference buffers. The buffers live until the
// process ends (planning may run on any thread at any time).
struct Selector {
  std::mutex mu;
  uint64_t gen = 0;    // 0: no weights
  uint64_t next = 1;
  std::vector<float> params, mean, inv_std;
  std::unordered_map<int, DevBuf> dev;
};

Selector& state() {
  static Selector* s = new Selector;
  return *s;
}

constexpr uint32_t kMagic = 0x4D54524C;  // "LRTM"

// Installs weights (params NULL: none): one network, or an ensemble of
// nparams / P networks that share the input normalisation and whose scores
// are averaged.
int install_weights(const float* params, int nparams, const float* mean, const float* stdv, int nfeat) {
  Selector& s = state();
  std::lock_guard<std::mutex> lock(s.mu);
  if (!params) {
    s.gen = 0;
    return LIBRETT_C_SUCCESS;
  }
  if (nparams <= 0 || nparams % mlp::P != 0 || nparams / mlp::P > kSelMaxModels || nfeat != mlp::F || !mean ||
      !stdv)
    return LIBRETT_C_INVALID_PARAMETER;
  for (int i = 0; i < nparams; i++)
    if (!std::isfinite(params[i])) return LIBRETT_C_INVALID_PARAMETER;
  s.params.assign(params, params + nparams);
  s.mean.assign(mean, mean + nfeat);
  s.inv_std.resize(nfeat);
  for (int i = 0; i < nfeat; i++) s.inv_std[i] = stdv[i] > 1e-6f ? 1.0f / stdv[i] : 1.0f;
  s.gen = s.next++;
  return LIBRETT_C_SUCCESS;
}

// Reads and installs a weights file.
int install_file(const char* path) {
  if (!path) return LIBRETT_C_INVALID_PARAMETER;
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return LIBRETT_C_INVALID_PARAMETER;
  uint32_t hdr[4] = {0, 0, 0, 0};
  uint32_t models = 1;
  std::vector<float> mean(mlp::F), stdv(mlp::F), params;
  bool ok = std::fread(hdr, sizeof(uint32_t), 4, fp) == 4 && hdr[0] == kMagic && (hdr[1] == 1 || hdr[1] == 2) &&
            (int)hdr[2] == mlp::F && (int)hdr[3] == mlp::H;
  // Version 2 adds the number of networks after the header.
  if (ok && hdr[1] == 2)
    ok = std::fread(&models, sizeof(uint32_t), 1, fp) == 1 && models >= 1 && (int)models <= kSelMaxModels;
  if (ok) params.resize((size_t)models * mlp::P);
  ok = ok && std::fread(mean.data(), sizeof(float), mlp::F, fp) == (size_t)mlp::F;
  ok = ok && std::fread(stdv.data(), sizeof(float), mlp::F, fp) == (size_t)mlp::F;
  ok = ok && std::fread(params.data(), sizeof(float), params.size(), fp) == params.size();
  std::fclose(fp);
  if (!ok) return LIBRETT_C_INVALID_PARAMETER;
  return install_weights(params.data(), (int)params.size(), mean.data(), stdv.data(), mlp::F);
}

// Directory of the shared library that holds this code.
std::string library_dir() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&library_dir), &info) == 0 || !info.dli_fname) return std::string();
  std::string p(info.dli_fname);
  const size_t slash = p.find_last_of('/');
  return slash == std::string::npos ? std::string(".") : p.substr(0, slash);
}

// On first use: the weights file named by LIBRETT_SELECTOR_WEIGHTS, or
// selector_weights.bin beside the library. LIBRETT_SELECTOR_WEIGHTS set to
// an empty string or "none" keeps the analytic model.
void load_from_env() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char* p = std::getenv("LIBRETT_SELECTOR_WEIGHTS");
    if (p) {
      if (!*p || std::strcmp(p, "none") == 0) return;
      if (install_file(p) != LIBRETT_C_SUCCESS)
        std::fprintf(stderr, "librett: cannot load selector weights from %s; using the analytic model\n", p);
      return;
    }
    const std::string def = library_dir() + "/selector_weights.bin";
    FILE* fp = std::fopen(def.c_str(), "rb");
    if (!fp) return;
    std::fclose(fp);
    if (install_file(def.c_str()) != LIBRETT_C_SUCCESS)
      std::fprintf(stderr, "librett: cannot load selector weights from %s; using the analytic model\n", def.c_str());
  });
}

// The run-time setters: the first-use load happens before them, so what
// they install replaces it instead of being replaced by it.
int set_weights(const float* params, int nparams, const float* mean, const float* stdv, int nfeat) {
  load_from_env();
  return install_weights(params, nparams, mean, stdv, nfeat);
}

int load_weights(const char* path) {
  load_from_env();
  return install_file(path);
}

}  // namespace

namespace {

int line_alignment(long long bytes) {
  int k = 0;
  while (k < 7 && bytes > 0 && (bytes & 1) == 0) {
    bytes >>= 1;
    k++;
  }
  return k;
}

}  // namespace

void selector_features(const Problem& full, const Problem& red, const Candidate& c, float* f) {
  for (int i = 0; i < mlp::F; i++) f[i] = 0.0f;
  f[0] = (float)red.rank;
  for (int i = 0; i < kSelMaxRank && i < red.rank; i++) {
    f[1 + i] = lg(red.dim[i]);
    f[8 + i] = (float)(red.perm[i] + 1);
  }
  f[15] = lg(full.elem);
  f[16] = lg((double)full.vol * full.elem);
  // How well the contiguous runs fill 128-byte lines: log2 of the largest
  // power of two (at most 128) dividing the byte length of the input's and
  // of the output's leading run, the smaller of the two (7: every row starts
  // on a line, as in power-of-two shapes; 2 or 3: rows straddle lines).
  f[17] = (float)std::min(line_alignment((long long)red.dim[0] * full.elem),
                          line_alignment((long long)red.dim[red.perm[0]] * full.elem));
  if (c.method >= kTrivial && c.method <= kTiledCopy) f[18 + (c.method - kTrivial)] = 1.0f;
  f[23] = lg(c.tile_x);
  f[24] = lg(c.tile_y);
  f[25] = lg(c.vec_x);
  f[26] = lg(c.vec_y);
  f[27] = lg(c.word);
  f[28] = lg(c.copy_tx);
  f[29] = (float)c.nreg;
  f[30] = lg(c.split.numSplit);
  f[31] = lg(c.split.volMmk);
  f[32] = lg(c.split.volMbar);
  f[33] = lg(c.split.volMm);
  f[34] = lg(c.split.volMk);
  f[35] = lg((double)c.grid_x * c.grid_y);
  f[36] = lg(c.block);
  f[37] = c.lds_bytes / 1024.0f;
  f[38] = (float)c.active_blocks_per_cu;
  const ModelStats& m = c.model;
  f[39] = lg(1 + m.read_lines);
  f[40] = lg(1 + m.write_lines);
  f[41] = lg(1 + m.partial_lines);
  f[42] = lg(1 + m.read_tran);
  f[43] = lg(1 + m.write_tran);
  f[44] = lg(1 + m.items);
  f[45] = lg(1 + m.steps_per_block);
  // Modelled times relative to the HBM roofline of the move (8 TB/s).
  const double t_roof = std::max(1.0, (double)c.bytes) / 8.0e12;
  f[46] = lg(m.predicted_s / t_roof);
  f[47] = lg(m.t_stream / t_roof);
  // Packed: work order and whether the plan is of the full (not the
  // rank-reduced) problem. Tiled: the extra dims in the plane's x and y
  // groups, and the tile order across XCDs; TiledCopy: its tile order.
  const bool packed = c.method == kPacked || c.method == kPackedSplit;
  f[48] = (float)(packed ? c.packed.order : c.method == kTiled ? (c.xgroup - 1) + (c.ygroup - 1) : 0);
  f[49] = packed ? (std::strcmp(c.source, "full") == 0 ? 1.0f : 0.0f) : (float)tile_order(c);
}

uint64_t selector_generation() {
  load_from_env();
  Selector& s = state();
  std::lock_guard<std::mutex> lock(s.mu);
  return s.gen;
}

double selector_margin() {
  static const double m = [] {
    const char* v = std::getenv("LIBRETT_SELECTOR_MARGIN");
    return (v && *v) ? std::atof(v) : kSelMargin;
  }();
  return m;
}

int selector_choose(const std::vector<Candidate>& cands, const Problem& full, int device) {
  load_from_env();
  Selector& s = state();
  std::lock_guard<std::mutex> lock(s.mu);
  if (!s.gen || cands.empty()) return -1;
  for (int i = 0; i < (int)cands.size(); i++)
    if (cands[i].method == kTrivial) return i;  // a plain copy is exclusive, as in the analytic choice
  const Problem red = reduce_ranks(full);
  if (red.rank > kSelMaxRank) return -1;
  const int n = (int)cands.size();
  std::vector<float> x((size_t)n * mlp::F);
  for (int i = 0; i < n; i++) {
    float* f = x.data() + (size_t)i * mlp::F;
    selector_features(full, red, cands[i], f);
    for (int k = 0; k < mlp::F; k++) {
      f[k] = (f[k] - s.mean[k]) * s.inv_std[k];
      if (!std::isfinite(f[k])) return -1;
    }
  }
  DevBuf& d = s.dev[device];
  bool ok = true;
  if (!d.stream) ok = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) == hipSuccess;
  const int models = (int)(s.params.size() / mlp::P);
  if (ok && !d.params) {
    ok = dev_malloc((void**)&d.params, (size_t)kSelMaxModels * mlp::P * sizeof(float)) == hipSuccess;
    if (!ok) d.params = nullptr;
    d.gen = 0;
  }
  if (ok && d.cap < n) {
    dev_free(d.x);
    dev_free(d.out);
    d.x = d.out = nullptr;
    d.cap = 0;
    const int cap = (n + 63) / 64 * 64;
    ok = dev_malloc((void**)&d.x, (size_t)cap * mlp::F * sizeof(float)) == hipSuccess &&
         dev_malloc((void**)&d.out, (size_t)kSelMaxModels * cap * sizeof(float)) == hipSuccess;
    if (ok) d.cap = cap;
  }
  if (ok && d.gen != s.gen) {
    ok = hipMemcpyAsync(d.params, s.params.data(), s.params.size() * sizeof(float), hipMemcpyHostToDevice,
                        d.stream) == hipSuccess;
    if (ok) d.gen = s.gen;
  }
  // One launch scores every candidate with every network (network k writes
  // out[k * n ..]); the score is their mean.
  std::vector<float> all((size_t)models * n), score(n, 0.0f);
  ok = ok &&
       hipMemcpyAsync(d.x, x.data(), x.size() * sizeof(float), hipMemcpyHostToDevice, d.stream) == hipSuccess &&
       mlp::launch_forward(d.params, d.x, n, d.out, d.stream, models) == hipSuccess &&
       hipMemcpyAsync(all.data(), d.out, all.size() * sizeof(float), hipMemcpyDeviceToHost, d.stream) ==
           hipSuccess &&
       hipStreamSynchronize(d.stream) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    d.gen = 0;
    return -1;
  }
  for (int k = 0; k < models; k++)
    for (int i = 0; i < n; i++) score[i] += all[(size_t)k * n + i];
  for (int i = 0; i < n; i++) score[i] /= (float)models;
  int best = -1;
  for (int i = 0; i < n; i++) {
    if (!std::isfinite(score[i])) return -1;
    if (best < 0 || score[i] < score[best]) best = i;
  }
  // The analytic model's choice stands unless the selector predicts its own
  // pick to be faster by more than the margin.
  const int mc = choose_by_model(cands);
  if (mc >= 0 && mc < n && score[mc] - score[best] < selector_margin()) return mc;
  return best;
}

void selector_prepare(int device) {
  load_from_env();
  Selector& s = state();
  std::lock_guard<std::mutex> lock(s.mu);
  if (!s.gen) return;
  DevBuf& d = s.dev[device];
  bool ok = true;
  if (!d.stream) ok = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) == hipSuccess;
  if (ok && !d.params) {
    ok = dev_malloc((void**)&d.params, (size_t)kSelMaxModels * mlp::P * sizeof(float)) == hipSuccess;
    if (!ok) d.params = nullptr;
    d.gen = 0;
  }
  if (ok && d.cap == 0) {
    dev_free(d.x);
    dev_free(d.out);
    d.x = d.out = nullptr;
    const int cap = 256;
    ok = dev_malloc((void**)&d.x, (size_t)cap * mlp::F * sizeof(float)) == hipSuccess &&
         dev_malloc((void**)&d.out, (size_t)kSelMaxModels * cap * sizeof(float)) == hipSuccess;
    if (ok) d.cap = cap;
  }
  if (ok && d.gen != s.gen) {
    ok = hipMemcpyAsync(d.params, s.params.data(), s.params.size() * sizeof(float), hipMemcpyHostToDevice,
                        d.stream) == hipSuccess &&
         hipStreamSynchronize(d.stream) == hipSuccess;
    if (ok) d.gen = s.gen;
  }
  // One inference on zeros, so that loading the kernel's code object is paid
  // here and not in the first plan.
  if (ok && !d.warm) {
    ok = hipMemsetAsync(d.x, 0, (size_t)mlp::F * sizeof(float), d.stream) == hipSuccess &&
         mlp::launch_forward(d.params, d.x, 1, d.out, d.stream, 1) == hipSuccess &&
         hipStreamSynchronize(d.stream) == hipSuccess;
    if (ok) d.warm = true;
  }
  if (!ok) (void)hipGetLastError();
}

void selector_release_buffers(const std::function<void()>& then) {
  Selector& s = state();
  std::lock_guard<std::mutex> lock(s.mu);
  for (auto& kv : s.dev) {
    DevBuf& d = kv.second;
    if (d.stream) (void)hipStreamSynchronize(d.stream);
    dev_free(d.params);
    dev_free(d.x);
    dev_free(d.out);
    d.params = d.x = d.out = nullptr;
    d.cap = 0;
    d.gen = 0;
  }
  (void)hipGetLastError();
  then();
}

}  // namespace lrt

using namespace lrt;

extern "C" {

void librett_c_mlp_sizes(int* nfeat, int* hidden, int* nparams) {
  if (nfeat) *nfeat = mlp::F;
  if (hidden) *hidden = mlp::H;
  if (nparams) *nparams = mlp::P;
}

int librett_c_selector_set(const float* params, int nparams, const float* mean, const float* stdv, int nfeat) {
  return set_weights(params, nparams, mean, stdv, nfeat);
}

int librett_c_selector_load(const char* path) { return load_weights(path); }

unsigned long long librett_c_selector_generation(void) { return selector_generation(); }

double librett_c_selector_margin(void) { return selector_margin(); }

int librett_c_mlp_forward(const float* params, const float* x, int n, float* out, void* stream) {
  if (!params || !x || !out || n < 0) return LIBRETT_C_INVALID_PARAMETER;
  if (mlp::launch_forward(params, x, n, out, (hipStream_t)stream) != hipSuccess) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  return LIBRETT_C_SUCCESS;
}

int librett_c_mlp_grad(const float* params, const float* x, const float* y, const float* w, const int* idx, int base,
                       int end, float inv_m, int nblocks, float* partial, float* loss, void* stream) {
  if (!params || !x || !y || !idx || !partial || !loss || nblocks < 1 || base < 0 || end < base)
    return LIBRETT_C_INVALID_PARAMETER;
  if (mlp::prepare_grad() != hipSuccess ||
      mlp::launch_grad(params, x, y, w, idx, base, end, inv_m, nblocks, partial, loss, (hipStream_t)stream) !=
          hipSuccess) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  return LIBRETT_C_SUCCESS;
}

int librett_c_mlp_adam(const float* partial, int nparts, float* params, float* m, float* v, float lr, float beta1,
                       float beta2, float c1, float c2, float eps, void* stream) {
  if (!partial || !params || !m || !v || nparts < 1) return LIBRETT_C_INVALID_PARAMETER;
  if (mlp::launch_adam(partial, nparts, params, m, v, lr, beta1, beta2, c1, c2, eps, (hipStream_t)stream) !=
      hipSuccess) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  return LIBRETT_C_SUCCESS;
}

int librett_c_mlp_train(const float* x, const float* y, const float* w, const float* w_host, int n, float* params,
                        int epochs, float lr, int minibatch, unsigned int seed, float* epoch_loss, void* stream_) {
  if (!x || !y || !params || n < 1 || epochs < 0 || minibatch < mlp::S) return LIBRETT_C_INVALID_PARAMETER;
  if ((w == nullptr) != (w_host == nullptr)) return LIBRETT_C_INVALID_PARAMETER;
  hipStream_t stream = (hipStream_t)stream_;
  const int nblocks = minibatch / mlp::S;
  const int mb = nblocks * mlp::S;
  int* d_idx = nullptr;
  float *d_part = nullptr, *d_m = nullptr, *d_v = nullptr, *d_loss = nullptr;
  bool ok = dev_malloc((void**)&d_idx, (size_t)n * sizeof(int)) == hipSuccess &&
            dev_malloc((void**)&d_part, (size_t)nblocks * mlp::P * sizeof(float)) == hipSuccess &&
            dev_malloc((void**)&d_m, mlp::P * sizeof(float)) == hipSuccess &&
            dev_malloc((void**)&d_v, mlp::P * sizeof(float)) == hipSuccess &&
            dev_malloc((void**)&d_loss, (size_t)nblocks * sizeof(float)) == hipSuccess;
  std::vector<float> zeros(mlp::P, 0.0f);
  ok = ok && hipMemcpyAsync(d_m, zeros.data(), mlp::P * sizeof(float), hipMemcpyHostToDevice, stream) == hipSuccess &&
       hipMemcpyAsync(d_v, zeros.data(), mlp::P * sizeof(float), hipMemcpyHostToDevice, stream) == hipSuccess &&
       mlp::prepare_grad() == hipSuccess;
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937 rng(seed);
  const double beta1 = 0.9, beta2 = 0.999;
  long long step = 0;
  std::vector<float> lossv(nblocks);
  for (int ep = 0; ok && ep < epochs; ep++) {
    std::shuffle(order.begin(), order.end(), rng);
    ok = hipMemcpyAsync(d_idx, order.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, stream) == hipSuccess &&
         hipStreamSynchronize(stream) == hipSuccess;
    double sse = 0, wsum = 0;
    for (int base = 0; ok && base < n; base += mb) {
      const int end = std::min(n, base + mb);
      step++;
      const float c1 = (float)(1.0 / (1.0 - std::pow(beta1, (double)step)));
      const float c2 = (float)(1.0 / (1.0 - std::pow(beta2, (double)step)));
      // The minibatch's weight sum normalises its loss.
      double m = end - base;
      if (w_host) {
        m = 0;
        for (int k = base; k < end; k++) m += w_host[order[k]];
      }
      wsum += m;
      ok = mlp::launch_grad(params, x, y, w, d_idx, base, end, (float)(1.0 / std::max(m, 1e-30)), nblocks, d_part,
                            d_loss, stream) == hipSuccess &&
           mlp::launch_adam(d_part, nblocks, params, d_m, d_v, lr, (float)beta1, (float)beta2, c1, c2, 1e-8f,
                            stream) == hipSuccess;
      if (ok && epoch_loss) {
        ok = hipMemcpyAsync(lossv.data(), d_loss, nblocks * sizeof(float), hipMemcpyDeviceToHost, stream) ==
                 hipSuccess &&
             hipStreamSynchronize(stream) == hipSuccess;
        for (float l : lossv) sse += l;
      }
    }
    if (epoch_loss) epoch_loss[ep] = (float)(sse / std::max(wsum, 1e-30));
  }
  ok = ok && hipStreamSynchronize(stream) == hipSuccess;
  if (d_idx) (void)dev_free(d_idx);
  if (d_part) (void)dev_free(d_part);
  if (d_m) (void)dev_free(d_m);
  if (d_v) (void)dev_free(d_v);
  if (d_loss) (void)dev_free(d_loss);
  if (!ok) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  return LIBRETT_C_SUCCESS;
}

void librett_c_mlp_host_forward(const float* params, const float* x, int n, float* out) {
  mlp::host_forward(params, x, n, out);
}

int librett_c_mlp_host_grad(const float* params, const float* x, const float* y, const float* w, const int* idx,
                            int count, float inv_m, float* grad, float* loss) {
  if (!params || !x || !y || !idx || !grad || !loss || count < 0 || count > mlp::S) return LIBRETT_C_INVALID_PARAMETER;
  mlp::host_grad(params, x, y, w, idx, count, inv_m, grad, loss);
  return LIBRETT_C_SUCCESS;
}

void librett_c_mlp_host_adam(const float* partial, int nparts, float* params, float* m, float* v, float lr,
                             float beta1, float beta2, float c1, float c2, float eps) {
  mlp::host_adam(partial, nparts, params, m, v, lr, beta1, beta2, c1, c2, eps);
}

}  // extern "C"
