// extern "C" entry points of libticc_hip.so: argument validation and dispatch.
#include <cmath>

#include <vector>

#include "ticc_common.h"

namespace ticc {
const char* last_error();
int release_workspace();
int contract_w(int N, int n_ang, int n_block, const int* block_offsets, const int* block_channels, const double* Br,
               const double* Bi, double norm, int nR, const double* R, const double* V, const double* U,
               const double* Eint, double mu, double* W, hipStream_t stream);
int ldmd_init(int nE, int N, const double* W0, const double* E_host, double mu, double* Y, hipStream_t stream);
int ldmd_init_capture(int nE, int N, const double* W0, const double* E_host, double mu, double* Y, hipStream_t stream);
int mlp_gemm(int mode, int M, int N, int K, const void* A, const void* B, const float* bias, const float* Z, float* Cf,
             void* Cb, void* CbT, hipStream_t stream);
int mlp_features(int n, int n_valid, const double* X4, double r_scale, void* F, void* FT, hipStream_t stream);
int mlp_grid_features(int n, int n_valid, int n_grid, const double* R, int n_coord, const double* coords, int ix, int iy,
                      int iphi, double r_scale, void* F, hipStream_t stream);
int mlp_loss(int n, int n_valid, const float* Y, const float* T, void* dZ, void* dZT, float* dbias, float* loss,
             hipStream_t stream);
int mlp_colsum(int M, int N, const float* X, float* out, hipStream_t stream);
int mlp_gemm_split(int mode, int M, int N, int K, const void* A, const void* A_lo, const void* B, const void* B_lo,
                   const float* bias, float* Cf, void* Cb, void* Cb_lo, void* CbT, hipStream_t stream);
int mlp_split_lo(int n, const float* x, void* lo, hipStream_t stream);
int mlp_radial_features(int n, int n_valid, const double* R, double r_scale, void* F, void* FT, hipStream_t stream);
int mlp_basis_loss(int n, int n_valid, int ld, const float* U, const int* gidx, const float* G, const float* T,
                   float* dZf, void* dZ, void* dZT, float* loss, hipStream_t stream);
int mlp_basis_eval(int nR, int n_grid, int nt, int ld, const float* U, const double* GT, const double* scale, double* V,
                   hipStream_t stream);
int mlp_adam(int rows, int cols, float* theta, const float* grad, float* m, float* v, float lr, float b1, float b2,
             float eps, int step, void* bf, void* bfT, hipStream_t stream);
int mlp_to_bf16(int rows, int cols, const float* x, void* bf, void* bfT, hipStream_t stream);

int get_smat(int nE, int N, const double* Y, const double* B, double Rmatch, const double* E_host, double mu,
             const double* Eint_host, const double* L_host, double* S_re, double* S_im, int* n_open,
             hipStream_t stream);
int profile_begin(int every, int max_pairs);
int mixed_stats(long long* inversions, long long* fallbacks, int reset);
int gj_set_walk(int fused_target, int update_target);
int profile_end(double* total_ms, double* union_ms, int* count, double* total_bytes, double* total_flop);
int bessel_eval(int n, const double* ell, const double* x, double* out, const double* nu, double* kout,
                hipStream_t stream);
}  // namespace ticc

using ticc::fail;

namespace ticc {
namespace {
// Scratch layout of one batched inversion: the padded matrices and their
// ping-pong partners, the Gauss-Jordan panels, row maps and pivot lists, and the
// row permutation and its inverse, each 256-byte aligned.
struct InvertLayout {
  size_t o0, o1, oP, oS, oV, oM, oI, total;
};
InvertLayout invert_layout(int n_batch, int Np) {
  InvertLayout l{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t n2 = (size_t)Np * Np;
  l.o0 = take(sizeof(double) * n_batch * n2);
  l.o1 = take(sizeof(double) * n_batch * n2);
  l.oP = take(sizeof(double) * GJ_POUT_PANELS * n_batch * Np * GJ_NB);
  l.oS = take(sizeof(int) * GJ_SRC_MAPS * n_batch * Np);
  l.oV = take(sizeof(int) * GJ_PIV_LISTS * n_batch * GJ_NB);
  l.oM = take(sizeof(int) * n_batch * Np);
  l.oI = take(sizeof(int) * n_batch * Np);
  l.total = off;
  return l;
}
}  // namespace

size_t invert_scratch_bytes(int n_batch, int Np) { return invert_layout(n_batch, Np).total; }

int invert_with_scratch(int n_batch, int n_channel, double* A, void* scratch, hipStream_t st) {
  const int Np = round_up(n_channel, GJ_NB);
  const InvertLayout l = invert_layout(n_batch, Np);
  char* w = static_cast<char*>(scratch);
  double* A0 = reinterpret_cast<double*>(w + l.o0);
  double* A1 = reinterpret_cast<double*>(w + l.o1);
  TICC_RETURN_IF(pad_identity(A, (long long)n_channel * n_channel, n_channel, A0, Np, n_batch, st));
  double* res = nullptr;
  int* pinv = reinterpret_cast<int*>(w + l.oI);
  TICC_RETURN_IF(gj_invert_raw(A0, A1, n_batch, Np, &res, reinterpret_cast<int*>(w + l.oM), pinv,
                               reinterpret_cast<double*>(w + l.oP), reinterpret_cast<int*>(w + l.oS),
                               reinterpret_cast<int*>(w + l.oV), st));
  double* other = (res == A0) ? A1 : A0;
  TICC_RETURN_IF(gj_unpermute(res, other, n_batch, Np, pinv, st));
  TICC_RETURN_IF(unpad(other, Np, A, (long long)n_channel * n_channel, n_channel, n_batch, st));
  TICC_CHECK_HIP(hipStreamSynchronize(st));
  return TICC_OK;
}
}  // namespace ticc

namespace {
hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

int check_channels(int n) {
  if (n < 1) return fail(TICC_ERR_VALUE, "n_channel must be positive, got %d", n);
  if (n > TICC_MAX_CHANNELS)
    return fail(TICC_ERR_UNSUPPORTED, "n_channel=%d exceeds the supported maximum %d", n, TICC_MAX_CHANNELS);
  return TICC_OK;
}
}  // namespace

extern "C" {

int ticc_abi_version(void) { return TICC_ABI_VERSION; }

const char* ticc_last_error(void) { return ticc::last_error(); }

int ticc_device_count(int* count) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  if (count) *count = n;
  return TICC_OK;
}

int ticc_set_device(int device) {
  int n = 0;
  ticc_device_count(&n);
  if (device < 0 || device >= n)
    return fail(TICC_ERR_VALUE, "requested HIP device %d, but %d device(s) are visible", device, n);
  TICC_CHECK_HIP(hipSetDevice(device));
  return TICC_OK;
}

int ticc_release_workspace(void) { return ticc::release_workspace(); }

int ticc_set_lane(int lane) { return ticc::set_lane(lane); }

int ticc_contract_w(int n_channel, int n_ang, int n_block, const int* block_offsets, const int* block_channels,
                    const double* B_real, const double* B_imag, double normalization, int n_R, const double* R,
                    const double* V, const double* U, const double* E_int, double reduced_mass, double* W,
                    void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  if (n_ang < 1 || n_block < 0 || n_R < 0) return fail(TICC_ERR_VALUE, "invalid contraction sizes");
  if (!block_offsets || (n_block > 0 && (!block_channels || !B_real)) || !U || !E_int || !W || (n_R > 0 && (!R || !V)))
    return fail(TICC_ERR_VALUE, "contraction: null pointer argument");
  if (block_offsets[0] != 0) return fail(TICC_ERR_VALUE, "block_offsets[0] must be 0");
  for (int k = 0; k < n_block; ++k)
    if (block_offsets[k + 1] < block_offsets[k]) return fail(TICC_ERR_VALUE, "block_offsets must be non-decreasing");
  if (block_offsets[n_block] > n_channel) return fail(TICC_ERR_VALUE, "basis rows exceed n_channel");
  if (!(reduced_mass > 0.0)) return fail(TICC_ERR_VALUE, "reduced_mass must be positive");
  return ticc::contract_w(n_channel, n_ang, n_block, block_offsets, block_channels, B_real, B_imag, normalization, n_R,
                          R, V, U, E_int, reduced_mass, W, as_stream(stream));
}

int ticc_ldmd_init(int n_energy, int n_channel, const double* W0, const double* energies, double reduced_mass,
                   double* Y, void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  if (n_energy < 1) return fail(TICC_ERR_VALUE, "At least one total energy is required");
  if (!W0 || !energies || !Y) return fail(TICC_ERR_VALUE, "ldmd_init: null pointer argument");
  return ticc::ldmd_init(n_energy, n_channel, W0, energies, reduced_mass, Y, as_stream(stream));
}

namespace {
int sweep_args(int n_energy, int n_sector, const double* energies, const double* half_steps, const double* W_start,
               const double* W_mid, const double* W_end, const double* Y) {
  if (n_energy < 1) return fail(TICC_ERR_VALUE, "At least one total energy is required");
  if (n_sector < 0) return fail(TICC_ERR_VALUE, "n_sector must be non-negative");
  if (!energies || !Y || (n_sector > 0 && (!half_steps || !W_start || !W_mid || !W_end)))
    return fail(TICC_ERR_VALUE, "ldmd_propagate: null pointer argument");
  for (int s = 0; s < n_sector; ++s)
    if (!(half_steps[s] > 0.0) || !std::isfinite(half_steps[s]))
      return fail(TICC_ERR_VALUE, "radial_half_step must be positive, but got radial_half_step=%.17g", half_steps[s]);
  for (int e = 0; e < n_energy; ++e)
    if (!std::isfinite(energies[e])) return fail(TICC_ERR_VALUE, "Total energies must be finite");
  return TICC_OK;
}
}  // namespace

int ticc_ldmd_init_capture(int n_energy, int n_channel, const double* W0, const double* energies,
                           double reduced_mass, double* Y, void* stream) {
  TICC_RETURN_IF(check_channels(2 * n_channel));
  if (n_energy < 1) return fail(TICC_ERR_VALUE, "At least one total energy is required");
  if (!W0 || !energies || !Y) return fail(TICC_ERR_VALUE, "ldmd_init_capture: null pointer argument");
  return ticc::ldmd_init_capture(n_energy, n_channel, W0, energies, reduced_mass, Y, as_stream(stream));
}

int ticc_ldmd_propagate_capture(int n_energy, int n_channel, const double* energies, double reduced_mass,
                                int n_sector, const double* half_steps, const double* W_start, const double* W_mid,
                                const double* W_end, long long w_sector_stride, double* Y, void* stream) {
  TICC_RETURN_IF(check_channels(2 * n_channel));
  TICC_RETURN_IF(sweep_args(n_energy, n_sector, energies, half_steps, W_start, W_mid, W_end, Y));
  if (n_sector == 0) return TICC_OK;
  return ticc::ldmd_propagate(n_energy, 2 * n_channel, energies, reduced_mass, n_sector, half_steps, W_start, W_mid,
                              W_end, w_sector_stride, Y, as_stream(stream), n_channel, 1, 0, -1, nullptr);
}

int ticc_ldmd_propagate(int n_energy, int n_channel, const double* energies, double reduced_mass, int n_sector,
                        const double* half_steps, const double* W_start, const double* W_mid, const double* W_end,
                        long long w_sector_stride, double* Y, void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  TICC_RETURN_IF(sweep_args(n_energy, n_sector, energies, half_steps, W_start, W_mid, W_end, Y));
  if (n_sector == 0) return TICC_OK;
  return ticc::ldmd_propagate(n_energy, n_channel, energies, reduced_mass, n_sector, half_steps, W_start, W_mid, W_end,
                              w_sector_stride, Y, as_stream(stream), 0, 1, 0, -1, nullptr);
}

int ticc_ldmd_propagate_mixed(int n_energy, int n_channel, const double* energies, double reduced_mass, int n_sector,
                              const double* half_steps, const double* W_start, const double* W_mid, const double* W_end,
                              long long w_sector_stride, int refine_steps, double* Y, void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  TICC_RETURN_IF(sweep_args(n_energy, n_sector, energies, half_steps, W_start, W_mid, W_end, Y));
  if (refine_steps < 0 || refine_steps > 3)
    return fail(TICC_ERR_VALUE, "refine_steps must lie in [0, 3], got %d", refine_steps);
  if (n_sector == 0) return TICC_OK;
  return ticc::ldmd_propagate(n_energy, n_channel, energies, reduced_mass, n_sector, half_steps, W_start, W_mid, W_end,
                              w_sector_stride, Y, as_stream(stream), 0, 1, 0, refine_steps, nullptr);
}

int ticc_ldmd_propagate_blocks(int n_block, int n_energy, int n_channel, const double* energies, double reduced_mass,
                               int n_sector, const double* half_steps, const double* W_start, const double* W_mid,
                               const double* W_end, long long w_sector_stride, long long w_block_stride, double* Y,
                               void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  if (n_block < 1) return fail(TICC_ERR_VALUE, "n_block must be positive");
  TICC_RETURN_IF(sweep_args(n_energy, n_sector, energies, half_steps, W_start, W_mid, W_end, Y));
  if ((long long)n_block * n_energy > (1LL << 20)) return fail(TICC_ERR_VALUE, "too many block-energy pairs");
  if (n_sector == 0) return TICC_OK;
  std::vector<double> all((size_t)n_block * n_energy);
  for (int b = 0; b < n_block; ++b)
    for (int e = 0; e < n_energy; ++e) all[(size_t)b * n_energy + e] = energies[e];
  return ticc::ldmd_propagate(n_block * n_energy, n_channel, all.data(), reduced_mass, n_sector, half_steps, W_start,
                              W_mid, W_end, w_sector_stride, Y, as_stream(stream), 0, n_block, w_block_stride, -1, nullptr);
}

int ticc_ldmd_propagate_blocks_async(int n_block, int n_energy, int n_channel, double reduced_mass, int n_sector,
                                     const double* half_steps, const double* energies_dev, const double* half_steps_dev,
                                     const double* W_start, const double* W_mid, const double* W_end,
                                     long long w_sector_stride, long long w_block_stride, const double* Y_in, double* Y,
                                     void* scratch, long long* scratch_bytes, void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  if (n_block < 1 || n_energy < 1 || n_sector < 1) return fail(TICC_ERR_VALUE, "n_block, n_energy, n_sector must be positive");
  if ((long long)n_block * n_energy > (1LL << 20)) return fail(TICC_ERR_VALUE, "too many block-energy pairs");
  if (!half_steps || !scratch_bytes) return fail(TICC_ERR_VALUE, "ldmd_propagate_blocks_async: null pointer argument");
  for (int s = 0; s < n_sector; ++s)
    if (!(half_steps[s] > 0.0) || !std::isfinite(half_steps[s]))
      return fail(TICC_ERR_VALUE, "radial_half_step must be positive, but got radial_half_step=%.17g", half_steps[s]);
  if (scratch && (!energies_dev || !half_steps_dev || !W_start || !W_mid || !W_end || !Y_in || !Y))
    return fail(TICC_ERR_VALUE, "ldmd_propagate_blocks_async: null pointer argument");
  if (*scratch_bytes < 0) return fail(TICC_ERR_VALUE, "scratch_bytes must be non-negative");
  ticc::SweepAsync as{energies_dev, half_steps_dev, Y_in, scratch, (size_t)*scratch_bytes};
  TICC_RETURN_IF(ticc::ldmd_propagate(n_block * n_energy, n_channel, nullptr, reduced_mass, n_sector, half_steps,
                                      W_start, W_mid, W_end, w_sector_stride, Y, as_stream(stream), 0, n_block,
                                      w_block_stride, -1, &as));
  *scratch_bytes = (long long)as.scratch_bytes;
  return TICC_OK;
}

namespace {
struct CapturedGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};
}  // namespace

int ticc_graph_begin(void* stream) {
  if (!stream) return fail(TICC_ERR_VALUE, "graph capture needs a non-default stream");
  TICC_CHECK_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeRelaxed));
  return TICC_OK;
}

int ticc_graph_end(void* stream, void** graph) {
  if (!stream || !graph) return fail(TICC_ERR_VALUE, "graph_end: null argument");
  CapturedGraph* g = new CapturedGraph();
  hipError_t e = hipStreamEndCapture(as_stream(stream), &g->graph);
  if (e != hipSuccess) {
    delete g;
    return fail(TICC_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    delete g;
    return fail(TICC_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  *graph = g;
  return TICC_OK;
}

int ticc_graph_launch(void* graph, void* stream) {
  if (!graph) return fail(TICC_ERR_VALUE, "graph_launch: null graph");
  TICC_CHECK_HIP(hipGraphLaunch(static_cast<CapturedGraph*>(graph)->exec, as_stream(stream)));
  return TICC_OK;
}

int ticc_graph_destroy(void* graph) {
  if (!graph) return TICC_OK;
  CapturedGraph* g = static_cast<CapturedGraph*>(graph);
  if (g->exec) TICC_CHECK_HIP(hipGraphExecDestroy(g->exec));
  if (g->graph) TICC_CHECK_HIP(hipGraphDestroy(g->graph));
  delete g;
  return TICC_OK;
}

int ticc_get_smat(int n_energy, int n_channel, const double* Y, const double* B, double Rmatch,
                  const double* energies, double reduced_mass, const double* E_int, const double* L, double* S_real,
                  double* S_imag, int* n_open, void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  if (n_energy < 1) return fail(TICC_ERR_VALUE, "At least one total energy is required");
  if (!Y || !energies || !E_int || !L || !S_real || !S_imag || !n_open)
    return fail(TICC_ERR_VALUE, "get_smat: null pointer argument");
  if (!(Rmatch > 0.0) || !(reduced_mass > 0.0))
    return fail(TICC_ERR_VALUE, "Rmatch and reduced_mass must be positive, but got Rmatch=%.17g, reduced_mass=%.17g",
                Rmatch, reduced_mass);
  for (int i = 0; i < n_channel; ++i)
    if (!(L[i] >= 0.0)) return fail(TICC_ERR_VALUE, "Orbital angular momenta L must be non-negative");
  return ticc::get_smat(n_energy, n_channel, Y, B, Rmatch, energies, reduced_mass, E_int, L, S_real, S_imag, n_open,
                        as_stream(stream));
}

int ticc_bessel_eval(int n, const double* ell, const double* x, double* out, const double* nu, double* kout,
                     void* stream) {
  if (n < 0) return fail(TICC_ERR_VALUE, "n must be non-negative");
  if (out && (!ell || !x)) return fail(TICC_ERR_VALUE, "bessel_eval: null pointer argument");
  if (kout && (!nu || !x)) return fail(TICC_ERR_VALUE, "bessel_eval: null pointer argument");
  TICC_RETURN_IF(ticc::bessel_eval(n, ell, x, out, nu, kout, as_stream(stream)));
  TICC_CHECK_HIP(hipStreamSynchronize(as_stream(stream)));
  return TICC_OK;
}

int ticc_invert_batched(int n_batch, int n_channel, double* A, void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  if (n_batch < 1 || !A) return fail(TICC_ERR_VALUE, "invert_batched: invalid arguments");
  const size_t bytes = ticc::invert_scratch_bytes(n_batch, ticc::round_up(n_channel, ticc::GJ_NB));
  void* base = nullptr;
  TICC_RETURN_IF(ticc::workspace(ticc::WS_INVERT, bytes, &base));
  return ticc::invert_with_scratch(n_batch, n_channel, A, base, as_stream(stream));
}

int ticc_mixed_stats(long long* inversions, long long* fallbacks, int reset) {
  if (!inversions || !fallbacks) return fail(TICC_ERR_VALUE, "mixed_stats: null pointer argument");
  return ticc::mixed_stats(inversions, fallbacks, reset);
}

int ticc_invert_batched_mixed(int n_batch, int n_channel, int refine_steps, double* A, void* stream) {
  TICC_RETURN_IF(check_channels(n_channel));
  if (n_batch < 1 || !A) return fail(TICC_ERR_VALUE, "invert_batched_mixed: invalid arguments");
  if (refine_steps < 0 || refine_steps > 3)
    return fail(TICC_ERR_VALUE, "refine_steps must lie in [0, 3], got %d", refine_steps);
  hipStream_t st = as_stream(stream);
  const int Np = ticc::round_up(n_channel, ticc::GJ_NB);
  const size_t n2 = (size_t)Np * Np;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t oA = take(sizeof(double) * n_batch * n2), oX = take(sizeof(double) * n_batch * n2);
  const size_t oF0 = take(sizeof(float) * n_batch * n2), oF1 = take(sizeof(float) * n_batch * n2);
  const size_t oT = take(sizeof(float) * n_batch * n2), oX0 = take(sizeof(float) * n_batch * n2),
               oX1 = take(sizeof(float) * n_batch * n2);
  const size_t oP = take(sizeof(float) * ticc::GJ_POUT_PANELS * n_batch * Np * ticc::GJ_NB);
  const size_t oS = take(sizeof(int) * ticc::GJ_SRC_MAPS * n_batch * Np),
               oV = take(sizeof(int) * ticc::GJ_PIV_LISTS * n_batch * ticc::GJ_NB);
  const size_t oM = take(sizeof(int) * n_batch * Np), oI = take(sizeof(int) * n_batch * Np);
  void* base = nullptr;
  TICC_RETURN_IF(ticc::workspace(ticc::WS_INVERT, off, &base));
  char* w = static_cast<char*>(base);
  auto D = [&](size_t o) { return reinterpret_cast<double*>(w + o); };
  auto F = [&](size_t o) { return reinterpret_cast<float*>(w + o); };
  auto I = [&](size_t o) { return reinterpret_cast<int*>(w + o); };
  TICC_RETURN_IF(ticc::pad_identity(A, (long long)n_channel * n_channel, n_channel, D(oA), Np, n_batch, st));
  TICC_RETURN_IF(ticc::to_f32(D(oA), F(oF0), (long long)n_batch * n2, st));
  float* r32 = nullptr;
  TICC_RETURN_IF(ticc::gj_invert_raw_f32(F(oF0), F(oF1), n_batch, Np, &r32, I(oM), I(oI), F(oP), I(oS), I(oV), st));
  TICC_RETURN_IF(ticc::gj_unpermute_f32(r32, D(oX), n_batch, Np, I(oI), st));
  if (refine_steps > 0) TICC_RETURN_IF(ticc::gj_unpermute_f32f32(r32, F(oX0), n_batch, Np, I(oI), st));
  float* X32[2] = {F(oX0), F(oX1)};
  for (int it = 0; it < refine_steps; ++it) {
    TICC_RETURN_IF(
        ticc::refine_gemm_f64(Np, n_batch, D(oA), D(oX), -1.0, nullptr, 0.0, 1.0, nullptr, F(oT), nullptr, st));
    TICC_RETURN_IF(ticc::refine_gemm_f32(Np, n_batch, X32[it & 1], F(oT), 1.0, D(oX), 1.0, 0.0, D(oX),
                                         it + 1 < refine_steps ? X32[(it + 1) & 1] : nullptr, st));
  }
  TICC_RETURN_IF(ticc::unpad(D(oX), Np, A, (long long)n_channel * n_channel, n_channel, n_batch, st));
  TICC_CHECK_HIP(hipStreamSynchronize(st));
  return TICC_OK;
}

int ticc_dgemm_batched(int m, int n, int k, double alpha, const double* A, long long strideA, const double* B,
                       long long strideB, double beta, double* C, long long strideC, int batch, void* stream) {
  if (!A || !B || !C) return fail(TICC_ERR_VALUE, "dgemm: null pointer argument");
  return ticc::dgemm_batched(false, false, m, n, k, alpha, A, k, strideA, B, n, strideB, beta, C, n, strideC, batch,
                             as_stream(stream));
}

int ticc_mlp_gemm(int mode, int M, int N, int K, const void* A, const void* B, const float* bias, const float* Z,
                  float* Cf, void* Cb, void* CbT, void* stream) {
  if (!A || !B) return fail(TICC_ERR_VALUE, "mlp_gemm: null operand");
  if ((mode == 1 || mode == 2) && !bias) return fail(TICC_ERR_VALUE, "mlp_gemm: bias required");
  if (mode == 3 && !Z) return fail(TICC_ERR_VALUE, "mlp_gemm: pre-activations required");
  if (!Cf || ((mode == 1 || mode == 3) && (!Cb || !CbT))) return fail(TICC_ERR_VALUE, "mlp_gemm: null output");
  return ticc::mlp_gemm(mode, M, N, K, A, B, bias, Z, Cf, Cb, CbT, as_stream(stream));
}

int ticc_mlp_features(int n, int n_valid, const double* X4, double r_scale, void* F, void* FT, void* stream) {
  if (n <= 0 || n_valid < 0 || n_valid > n || !F || (n_valid > 0 && !X4))
    return fail(TICC_ERR_VALUE, "mlp_features: invalid arguments");
  return ticc::mlp_features(n, n_valid, X4, r_scale, F, FT, as_stream(stream));
}

int ticc_mlp_grid_features(int n, int n_valid, int n_grid, const double* R, int n_coord, const double* coords, int ix,
                           int iy, int iphi, double r_scale, void* F, void* stream) {
  if (n <= 0 || n_valid < 0 || n_valid > n || n_grid <= 0 || !F || !R || !coords || ix < 0 || iy < 0 || iphi < 0 ||
      ix >= n_coord || iy >= n_coord || iphi >= n_coord)
    return fail(TICC_ERR_VALUE, "mlp_grid_features: invalid arguments");
  return ticc::mlp_grid_features(n, n_valid, n_grid, R, n_coord, coords, ix, iy, iphi, r_scale, F, as_stream(stream));
}

int ticc_mlp_loss(int n, int n_valid, const float* Y, const float* T, void* dZ, void* dZT, float* dbias, float* loss,
                  void* stream) {
  if (n <= 0 || n_valid <= 0 || n_valid > n || !Y || !T || !dZ || !dZT || !dbias || !loss)
    return fail(TICC_ERR_VALUE, "mlp_loss: invalid arguments");
  return ticc::mlp_loss(n, n_valid, Y, T, dZ, dZT, dbias, loss, as_stream(stream));
}

int ticc_mlp_gemm_split(int mode, int M, int N, int K, const void* A, const void* A_lo, const void* B,
                        const void* B_lo, const float* bias, float* Cf, void* Cb, void* Cb_lo, void* CbT, void* stream) {
  if (!A || !B || !B_lo || !bias || !Cf) return fail(TICC_ERR_VALUE, "mlp_gemm_split: null operand");
  if (mode == 1 && (!Cb || !Cb_lo || !CbT)) return fail(TICC_ERR_VALUE, "mlp_gemm_split: null activation output");
  return ticc::mlp_gemm_split(mode, M, N, K, A, A_lo, B, B_lo, bias, Cf, Cb, Cb_lo, CbT, as_stream(stream));
}

int ticc_mlp_split_lo(int n, const float* x, void* lo, void* stream) {
  if (n <= 0 || !x || !lo) return fail(TICC_ERR_VALUE, "mlp_split_lo: invalid arguments");
  return ticc::mlp_split_lo(n, x, lo, as_stream(stream));
}

int ticc_mlp_radial_features(int n, int n_valid, const double* R, double r_scale, void* F, void* FT, void* stream) {
  if (n <= 0 || n_valid < 0 || n_valid > n || !F || (n_valid > 0 && !R))
    return fail(TICC_ERR_VALUE, "mlp_radial_features: invalid arguments");
  return ticc::mlp_radial_features(n, n_valid, R, r_scale, F, FT, as_stream(stream));
}

int ticc_mlp_basis_loss(int n, int n_valid, int ld, const float* U, const int* gidx, const float* G, const float* T,
                        float* dZf, void* dZ, void* dZT, float* loss, void* stream) {
  if (n <= 0 || n_valid <= 0 || n_valid > n || ld <= 0 || ld % 64 || !U || !gidx || !G || !T || !dZf || !dZ || !dZT ||
      !loss)
    return fail(TICC_ERR_VALUE, "mlp_basis_loss: invalid arguments");
  return ticc::mlp_basis_loss(n, n_valid, ld, U, gidx, G, T, dZf, dZ, dZT, loss, as_stream(stream));
}

int ticc_mlp_basis_eval(int nR, int n_grid, int nt, int ld, const float* U, const double* GT, const double* scale,
                        double* V, void* stream) {
  if (nR <= 0 || n_grid <= 0 || nt <= 0 || nt > 512 || nt > ld || !U || !GT || !scale || !V)
    return fail(TICC_ERR_VALUE, "mlp_basis_eval: invalid arguments");
  return ticc::mlp_basis_eval(nR, n_grid, nt, ld, U, GT, scale, V, as_stream(stream));
}

int ticc_mlp_colsum(int M, int N, const float* X, float* out, void* stream) {
  if (M <= 0 || N <= 0 || !X || !out) return fail(TICC_ERR_VALUE, "mlp_colsum: invalid arguments");
  return ticc::mlp_colsum(M, N, X, out, as_stream(stream));
}

int ticc_mlp_adam(int rows, int cols, float* theta, const float* grad, float* m, float* v, float lr, float beta1,
                  float beta2, float eps, int step, void* bf, void* bfT, void* stream) {
  if (rows <= 0 || cols <= 0 || step < 1 || !theta || !grad || !m || !v)
    return fail(TICC_ERR_VALUE, "mlp_adam: invalid arguments");
  return ticc::mlp_adam(rows, cols, theta, grad, m, v, lr, beta1, beta2, eps, step, bf, bfT, as_stream(stream));
}

int ticc_mlp_to_bf16(int rows, int cols, const float* x, void* bf, void* bfT, void* stream) {
  if (rows <= 0 || cols <= 0 || !x) return fail(TICC_ERR_VALUE, "mlp_to_bf16: invalid arguments");
  return ticc::mlp_to_bf16(rows, cols, x, bf, bfT, as_stream(stream));
}

int ticc_set_sweep_walk(int fused_target, int update_target) {
  return ticc::gj_set_walk(fused_target, update_target);
}

int ticc_profile_update_begin(int sample_every, int max_samples) {
  if (sample_every < 1 || max_samples < 1 || max_samples > (1 << 20))
    return fail(TICC_ERR_VALUE, "profile: sample_every and max_samples must be positive");
  return ticc::profile_begin(sample_every, max_samples);
}

int ticc_profile_update_end(double* total_ms, double* union_ms, int* count, double* total_bytes, double* total_flop) {
  if (!total_ms || !union_ms || !count || !total_bytes || !total_flop)
    return fail(TICC_ERR_VALUE, "profile: null pointer argument");
  return ticc::profile_end(total_ms, union_ms, count, total_bytes, total_flop);
}

}  // extern "C"
