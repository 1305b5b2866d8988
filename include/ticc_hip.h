/*
 * ticc_hip.h -- C ABI of the MI355X (gfx950) coupled-channel backend.
 *
 * One shared library, libticc_hip.so, built with hipcc for gfx950.  Every entry
 * point takes plain pointers and sizes; "device" pointers are HIP device
 * addresses, "host" pointers are ordinary CPU memory.  All work is enqueued on
 * `stream` (a hipStream_t passed as void*; NULL means the legacy default
 * stream) and is asynchronous unless stated otherwise.  Matrices are FP64,
 * row-major, contiguous.
 *
 * Return value: TICC_OK (0) or one of the TICC_ERR_* codes; ticc_last_error()
 * then returns a thread-local message.  The Python layer maps
 *   TICC_ERR_VALUE -> ValueError, TICC_ERR_UNSUPPORTED -> NotImplementedError,
 *   TICC_ERR_NONFINITE -> FloatingPointError, everything else -> RuntimeError,
 * which mirrors the reference's validate / logger.error / raise convention.
 *
 * The reference (PyTICC) has no foreign-function boundary: its hot path is
 * jax.numpy lowered by XLA.  Each entry point below replaces one of its
 * internal operator seams; the reference file:line is cited per function.
 */
#ifndef TICC_HIP_H
#define TICC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define TICC_ABI_VERSION 1

#define TICC_OK 0
#define TICC_ERR_VALUE 1
#define TICC_ERR_HIP 2
#define TICC_ERR_NOMEM 3
#define TICC_ERR_NONFINITE 4
#define TICC_ERR_UNSUPPORTED 5

/* Largest channel count of one propagated block. */
#define TICC_MAX_CHANNELS 1536

/* ABI version of the loaded library (TICC_ABI_VERSION). */
int ticc_abi_version(void);

/* Thread-local message describing the last non-zero status of this thread. */
const char* ticc_last_error(void);

/* Number of visible HIP devices (0 when no GPU is present).  Never fails. */
int ticc_device_count(int* count);

/* Make `device` current for the calling thread (hipSetDevice).
 * Replaces resolve_device / jax.devices (src/pyticc/propagation/device.py:52-71). */
int ticc_set_device(int device);

/* Free every library-owned workspace on the current device (synchronises it). */
int ticc_release_workspace(void);

/* Number of lanes of library state per device (see ticc_set_lane). */
#define TICC_LANES 4

/*
 * Select the calling thread's lane (no counterpart in the reference; for
 * concurrent sweeps from several host threads, as the bucketed batching
 * driver runs them).  The library's workspaces, auxiliary streams and events
 * are kept per device and lane: calls made on different lanes, from different
 * threads on different streams, never share them and may run concurrently.
 * Calls on one lane must not overlap.  Every thread starts on lane 0.
 */
int ticc_set_lane(int lane);

/*
 * W(R) contraction + radial-matrix assembly, kernel A.
 * Replaces contract_device / _contract_block_device
 * (src/pyticc/matrix/interaction/common.py:66-73, 153-211) and the device W
 * assembly W = U/R^2 + 2 mu V + 2 mu diag(E_int) (src/pyticc/propagation/runner.py:265-271).
 *
 *   n_channel        N
 *   n_ang            angular grid points per R (prod of the PES grid shape)
 *   n_block          number of helicity (K) blocks
 *   block_offsets    host int[n_block+1]: rows of block k are basis rows
 *                    [block_offsets[k], block_offsets[k+1])
 *   block_channels   device int[N]: channel index of each basis row
 *   B_real, B_imag   device (N, n_ang) weighted basis rows grouped by block;
 *                    B_imag may be NULL (atom-diatom)
 *   normalization    1/pi for AB+CD, 1 for A+BC
 *   n_R, R           device double[n_R] radial points
 *   V                device (n_R, n_ang) raw PES values at those points
 *   U, E_int         device (N, N) centrifugal matrix, device double[N] thresholds
 *   W                device (n_R, N, N) output, fully written
 */
int ticc_contract_w(int n_channel, int n_ang, int n_block, const int* block_offsets, const int* block_channels,
                    const double* B_real, const double* B_imag, double normalization, int n_R, const double* R,
                    const double* V, const double* U, const double* E_int, double reduced_mass, double* W,
                    void* stream);

/*
 * Initial log-derivative for the inelastic boundary condition,
 * Y_e = diag(sqrt|W0_ii - 2 mu E_e|).
 * Replaces initialize_logD_inelastic (src/pyticc/propagation/logd.py:48-65) as
 * called at src/pyticc/propagation/runner.py:284-291.
 *   W0 device (N, N) energy-independent W_base at the first radial point;
 *   energies host double[n_energy]; Y device (n_energy, N, N) output.
 */
int ticc_ldmd_init(int n_energy, int n_channel, const double* W0, const double* energies, double reduced_mass,
                   double* Y, void* stream);

/*
 * LDMD sweep of an energy batch over n_sector sectors, kernel B.
 * Replaces propagate_logD / _propagate_logD_jax
 * (src/pyticc/propagation/logd.py:288-330, 340-398), including the
 * energy-independent corrections (_correction_matrices, logd.py:143-164) and the
 * per-energy two-half-step update (_propagate_with_corrections, logd.py:168-194).
 *
 *   energies       host double[n_energy]
 *   half_steps     host double[n_sector], each > 0
 *   W_start, W_mid, W_end
 *                  device pointers to the energy-independent W_base of sector
 *                  0; sector s uses pointer + s * w_sector_stride (elements).
 *                  (W_points of 2 n_sector + 1 radial points: W_start = W,
 *                  W_mid = W + N*N, W_end = W + 2*N*N, stride 2*N*N.)
 *   Y              device (n_energy, N, N), in: Y at the first sector start,
 *                  out: Y at the last sector end.
 * A non-finite result returns TICC_ERR_NONFINITE (the reference silently
 * propagates NaN).  Synchronises `stream` once at the end for that check.
 */
int ticc_ldmd_propagate(int n_energy, int n_channel, const double* energies, double reduced_mass, int n_sector,
                        const double* half_steps, const double* W_start, const double* W_mid, const double* W_end,
                        long long w_sector_stride, double* Y, void* stream);

/*
 * LDMD sweep of several channel blocks in one batch: the CS / NNCC K blocks of
 * one (J, parity) problem (src/pyticc/propagation/runner.py:326-374, which
 * propagates all blocks through each radial window together), or any blocks of
 * one common size.  Each of the n_block blocks has its own W_base at
 * W_x + b * w_block_stride (sector s at + s * w_sector_stride) and is swept at
 * the same n_energy energies; Y is (n_block * n_energy, N, N), block-major.
 * Blocks smaller than N are padded by the caller with decoupled closed channels
 * (zero coupling, a large positive diagonal), which leaves their own channels'
 * sweep unchanged.
 */
int ticc_ldmd_propagate_blocks(int n_block, int n_energy, int n_channel, const double* energies, double reduced_mass,
                               int n_sector, const double* half_steps, const double* W_start, const double* W_mid,
                               const double* W_end, long long w_sector_stride, long long w_block_stride, double* Y,
                               void* stream);

/*
 * Mixed-precision LDMD sweep (the precision form; no counterpart in the
 * reference, whose sweep is jnp FP64 throughout, src/pyticc/propagation/logd.py:168-194).
 * Same arguments and results as ticc_ldmd_propagate, but each sector's two
 * inversions run as an FP32 Gauss-Jordan inversion (FP32 panels and FP32 MFMA
 * updates, the same pivot rule) followed by refine_steps (0..3) Newton steps
 * X <- X + X (I - A X): the residual I - A X from the FP64 A on the FP64 MFMA,
 * the correction on the FP32 MFMA.  Reference values, corrections, Q_mid, the
 * form steps and Y stay FP64.  A matrix whose first residual max|I - A X|
 * exceeds 1e-3 (TICC_MIXED_FALLBACK_TAU) is inverted in FP64 instead, so
 * ill-conditioned sector matrices near log-derivative poles keep FP64 accuracy.
 * Two steps keep the S matrix within 1e-6 of the FP64 sweep on the benchmark
 * systems.
 */
int ticc_ldmd_propagate_mixed(int n_energy, int n_channel, const double* energies, double reduced_mass, int n_sector,
                              const double* half_steps, const double* W_start, const double* W_mid, const double* W_end,
                              long long w_sector_stride, int refine_steps, double* Y, void* stream);

/*
 * Counters of the calling thread's mixed-precision sweeps since the last reset:
 * the sector-matrix inversions done, and how many of them were redone in FP64
 * because the FP32 inverse's residual max|I - A X| exceeded 1e-3 (near a
 * log-derivative pole the sector matrix is too ill-conditioned for FP32
 * refinement to reach FP64 accuracy).  reset != 0 zeroes them after reading.
 */
int ticc_mixed_stats(long long* inversions, long long* fallbacks, int reset);

/*
 * Enqueue-only form of ticc_ldmd_propagate_blocks, for capture into a HIP graph
 * (the batching form: many blocks' sweeps replayed as one graph launch; the
 * reference's batched block loop is src/pyticc/propagation/runner.py:326-374).
 * It performs no host-device copy, memset, allocation, host synchronisation or
 * finite check:
 *   energies_dev    device double[n_block * n_energy], block-major
 *   half_steps      host double[n_sector] (its values are baked into the kernels'
 *                   arguments), half_steps_dev the same values on the device
 *   Y_in            device initial Y (n_block * n_energy, N, N); Y receives the
 *                   result, so a replayed graph restarts from Y_in every time
 *   scratch         caller-owned device memory of *scratch_bytes bytes; with
 *                   scratch == NULL only *scratch_bytes is set (the size needed)
 *                   and nothing is enqueued.
 * FP64 inelastic sweep; the caller checks Y for non-finite values.
 */
int ticc_ldmd_propagate_blocks_async(int n_block, int n_energy, int n_channel, double reduced_mass, int n_sector,
                                     const double* half_steps, const double* energies_dev, const double* half_steps_dev,
                                     const double* W_start, const double* W_mid, const double* W_end,
                                     long long w_sector_stride, long long w_block_stride, const double* Y_in, double* Y,
                                     void* scratch, long long* scratch_bytes, void* stream);

/*
 * Stream capture into a HIP graph (no counterpart in the reference): begin
 * (relaxed capture mode) on a non-default stream, enqueue work (e.g. the sweeps
 * above), end -> an instantiated graph handle; launch replays it on a stream;
 * destroy frees it.  Work captured must not copy from pageable host memory or
 * allocate; device buffers it uses must outlive the graph.
 */
int ticc_graph_begin(void* stream);
int ticc_graph_end(void* stream, void** graph);
int ticc_graph_launch(void* graph, void* stream);
int ticc_graph_destroy(void* graph);

/*
 * Capture boundary condition (Propagation(mode="capture")): complex symmetric
 * log derivatives Z carried in the real form [[Re Z, -Im Z], [Im Z, Re Z]]
 * (2N x 2N), on which sums, products, inverses and real diagonal shifts act as
 * on Z, so the sweep is the real sweep of size 2N.
 *
 * Replaces initialize_logD_capture (src/pyticc/propagation/logd.py:72-91), as
 * called at src/pyticc/propagation/runner.py:284-291: Z_e = diag(z),
 * z_i = sqrt(d) for d = W0_ii - 2 mu E_e >= 0, -i sqrt(-d) otherwise.
 *   W0 device (N, N); Y device (n_energy, 2N, 2N) output (real form).
 */
int ticc_ldmd_init_capture(int n_energy, int n_channel, const double* W0, const double* energies,
                           double reduced_mass, double* Y, void* stream);

/*
 * LDMD sweep of complex log derivatives in real form (the capture mode of
 * propagate_logD, src/pyticc/propagation/logd.py:288-330, complex128 there).
 * Same arguments as ticc_ldmd_propagate with n_channel = N, but W_start,
 * W_mid, W_end point to the real form of W_base, block-diag(W, W) (2N x 2N),
 * and Y holds the real forms of the energies' Z (n_energy, 2N, 2N).  The
 * end-of-sector symmetrisation is Z -> (Z + Z^T)/2 in real form.
 */
int ticc_ldmd_propagate_capture(int n_energy, int n_channel, const double* energies, double reduced_mass,
                                int n_sector, const double* half_steps, const double* W_start, const double* W_mid,
                                const double* W_end, long long w_sector_stride, double* Y, void* stream);

/*
 * Asymptotic matching, kernel C: Y_SF = B^T Y B, then per energy the
 * closed-channel elimination and K -> S.
 * Replaces transform_logD_BF_to_SF (src/pyticc/match/asymptotic.py:82-110) and
 * get_Smat / _open_reaction_matrix / _reference_functions
 * (src/pyticc/match/smatrix.py:10-223), with the Riccati-Bessel and K_nu
 * log-derivative references of src/pyticc/match/bessel.py:7-45, 90-122
 * evaluated on the device.
 *
 *   Y        device (n_energy, N, N) BF log derivatives
 *   B        device (N, N) BF->SF orthogonal transform (NULL: identity)
 *   energies host double[n_energy]
 *   E_int, L host double[N]
 *   S_real, S_imag  device (n_energy, N, N) outputs; for energy e the leading
 *            n_open(e) x n_open(e) block holds S in flatnonzero(E_int < E)
 *            order, the rest is zero.
 *   n_open   host int[n_energy] output.
 * Raises TICC_ERR_VALUE if an energy equals a threshold or L < 0.
 */
int ticc_get_smat(int n_energy, int n_channel, const double* Y, const double* B, double Rmatch,
                  const double* energies, double reduced_mass, const double* E_int, const double* L, double* S_real,
                  double* S_imag, int* n_open, void* stream);

/*
 * Device reference functions (for tests): for each of n points,
 * out[4*i .. 4*i+3] = riccati j, n, j', n' of order ell[i] at x[i]
 * (src/pyticc/match/bessel.py:7-45);  kout[i] = K'_nu/K_nu at (nu[i], x[i])
 * (bessel.py:90-122).  Either output may be NULL.  All pointers device.
 */
int ticc_bessel_eval(int n, const double* ell, const double* x, double* out, const double* nu, double* kout,
                     void* stream);

/*
 * Batched in-place inverse of n_batch matrices (n_batch, N, N), device, by
 * blocked Gauss-Jordan with partial pivoting (the arithmetic of
 * jnp.linalg.solve(A, I), src/pyticc/propagation/logd.py:162).
 */
int ticc_invert_batched(int n_batch, int n_channel, double* A, void* stream);

/* Size of the fixed scratch array of ticc_invert_batched_noalloc, per GPU. */
#define TICC_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)

/*
 * ticc_invert_batched without any allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory -- code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * arithmetic and results as ticc_invert_batched, but its scratch (padded
 * copies, ping-pong partners, panels and row maps) lives in a fixed device array
 * of TICC_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a call
 * whose scratch would exceed it (about 12 x Np^2 bytes per matrix plus panels,
 * Np = N rounded up to 32) returns TICC_ERR_VALUE.  Calls that overlap on one GPU
 * share that array and race; calls on different GPUs do not.  Synchronises
 * `stream`.
 */
int ticc_invert_batched_noalloc(int n_batch, int n_channel, double* A, void* stream);

/*
 * Batched in-place inverse by the mixed-precision scheme of
 * ticc_ldmd_propagate_mixed: FP32 Gauss-Jordan with partial pivoting, then
 * refine_steps (0..3) Newton steps with the residual from the FP64 input.
 * (Test and utility entry; no counterpart in the reference.)
 */
int ticc_invert_batched_mixed(int n_batch, int n_channel, int refine_steps, double* A, void* stream);

/* Batched C = alpha A B + beta C for (m, k) x (k, n) row-major device matrices
 * with batch strides (elements); stride 0 broadcasts.  (Test and utility entry.) */
int ticc_dgemm_batched(int m, int n, int k, double alpha, const double* A, long long strideA, const double* B,
                       long long strideB, double beta, double* C, long long strideC, int batch, void* stream);

/*
 * Learned PES surrogate (no counterpart in the reference; the PES callback it
 * stands in for is the reference's interaction routine, src/pyticc/pes/adiabatic.py:17-103).
 * A 5 -> 256 x 4 -> 1 SiLU network in BF16 with FP32 accumulation; all
 * matrices row-major, BF16 buffers hold raw bfloat16 bits.
 *
 * C[M][N] = A[M][K] . B[N][K]^T with epilogue `mode`:
 *   0  Cf = C;
 *   1  Z = C + bias[col] -> Cf, a = SiLU(Z) -> Cb (M x N) and CbT (N x M), BF16;
 *   2  Cf = C + bias[col];
 *   3  dZ = C * SiLU'(Z[row][col]) -> Cf, Cb and CbT (BF16).
 * M, N multiples of 64, K a multiple of 16.
 */
int ticc_mlp_gemm(int mode, int M, int N, int K, const void* A, const void* B, const float* bias, const float* Z,
                  float* Cf, void* Cb, void* CbT, void* stream);

/* Features [r_scale/R, cos th_X, cos th_Y, cos phi, sin th_X sin th_Y cos phi] of
 * the first n_valid rows of X4 (n_valid x 4 doubles: R, th_X, th_Y, phi) into F
 * (n x 64 BF16, columns 5.. zero) and optionally FT (64 x n); rows past n_valid
 * get a fixed valid geometry. */
int ticc_mlp_features(int n, int n_valid, const double* X4, double r_scale, void* F, void* FT, void* stream);

/* The same features for sample i = (R[i / n_grid], grid point i % n_grid) of a
 * PES grid; coords column-major (n_coord x n_grid), angles in rows ix, iy, iphi. */
int ticc_mlp_grid_features(int n, int n_valid, int n_grid, const double* R, int n_coord, const double* coords, int ix,
                           int iy, int iphi, double r_scale, void* F, void* stream);

/* Mean-square loss of column 0 of Y (n x 64) against T over n_valid rows:
 * *loss += mean d^2, dZ / dZT column / row 0 = 2 d / n_valid (BF16), *dbias += sum of it. */
int ticc_mlp_loss(int n, int n_valid, const float* Y, const float* T, void* dZ, void* dZT, float* dbias, float* loss,
                  void* stream);

/*
 * Split-BF16 forward products of the surrogate: A and B each given as BF16 high
 * and low parts (x_lo = bf16(x - x_hi); A_lo may be NULL), C = A_hi B_hi^T +
 * A_hi B_lo^T + A_lo B_hi^T with FP32 accumulation (about 16 significant bits).
 * mode 1: Z = C + bias -> Cf, a = SiLU(Z) -> Cb (high), Cb_lo (low), CbT (high,
 * transposed); mode 2: Cf = C + bias.  Shapes as ticc_mlp_gemm.
 * split_lo: lo[i] = bf16(x[i] - bf16(x[i])) of an FP32 tensor of n elements.
 */
int ticc_mlp_gemm_split(int mode, int M, int N, int K, const void* A, const void* A_lo, const void* B,
                        const void* B_lo, const float* bias, float* Cf, void* Cb, void* Cb_lo, void* CbT, void* stream);
int ticc_mlp_split_lo(int n, const float* x, void* lo, void* stream);

/*
 * The angular-basis form of the surrogate: the network maps radial features to
 * the coefficients u_t(R) of fixed angular functions G_t (the bispherical terms of
 * the PES expansion), V = v_scale s(R) sum_t u_t(R) G_t.
 * radial_features: row i of F (n x 64 BF16, FT 64 x n optional) = [x, x^2, x^3,
 *   256 (x - bf16(x)), 256 (x^2 - bf16(x^2)), 256 (x^3 - bf16(x^3))], x = r_scale / R[i]
 *   (R resolved to about 16 bits; padding rows past n_valid: x = 1/2).
 * basis_loss: per row, pred = U[i][:] . G[gidx[i]][:] (ld columns, FP32), d = pred - T[i];
 *   *loss += sum d^2 / n_valid; dZf (FP32), dZ and dZT (BF16) = 2 d G[gidx[i]][:] / n_valid
 *   (zero rows past n_valid).  ld a multiple of 64.
 * basis_eval: V[r][g] = scale[r] sum_{t < nt} U[r][t] GT[t][g] (U FP32 with row stride ld,
 *   GT FP64 nt x n_grid, V FP64 nR x n_grid), nt <= 512.
 */
int ticc_mlp_radial_features(int n, int n_valid, const double* R, double r_scale, void* F, void* FT, void* stream);
int ticc_mlp_basis_loss(int n, int n_valid, int ld, const float* U, const int* gidx, const float* G, const float* T,
                        float* dZf, void* dZ, void* dZT, float* loss, void* stream);
int ticc_mlp_basis_eval(int nR, int n_grid, int nt, int ld, const float* U, const double* GT, const double* scale,
                        double* V, void* stream);

/* out[c] += sum_r X[r][c], X (M x N), N a multiple of 64. */
int ticc_mlp_colsum(int M, int N, const float* X, float* out, void* stream);

/* One Adam step (bias-corrected, step >= 1) of an FP32 rows x cols tensor; writes
 * the BF16 copy bf and the BF16 transpose bfT when they are not NULL. */
int ticc_mlp_adam(int rows, int cols, float* theta, const float* grad, float* m, float* v, float lr, float beta1,
                  float beta2, float eps, int step, void* bf, void* bfT, void* stream);

/* BF16 copy and transpose of an FP32 rows x cols tensor. */
int ticc_mlp_to_bf16(int rows, int cols, const float* x, void* bf, void* bfT, void* stream);

/*
 * Tile walk of the Gauss-Jordan update launches (no counterpart in the
 * reference; set by the sweep autotuner per size bucket): the number of
 * workgroups the column tiles of one rank-64 update are spread over, for the
 * update that runs beside the next panel chain (default 512) and for the
 * standalone updates (default 2048).  Applies to the calling thread's later
 * sweeps and inversions; 0 restores a default.  Results do not depend on it.
 */
int ticc_set_sweep_walk(int fused_target, int update_target);

/*
 * Instrumentation (no counterpart in the reference): live HIP-event timing of
 * the sweep's dominant kernel, the fused Gauss-Jordan launch (one pair's
 * rank-64 update beside the next pair's panel chain).  Between begin and end,
 * every sample_every-th batched inversion has each of its fused launches
 * bracketed by an event pair on the stream it runs on (at most max_samples
 * pairs).  end synchronises the device and returns the summed kernel time,
 * the length of the union of the timed launches' intervals on the device
 * clock (the time in which at least one of them ran: the sum when they never
 * overlap, less when launches of energy groups on separate streams run side by
 * side), the number of timed launches and their summed algorithmic bytes and
 * FP64 flop.
 */
int ticc_profile_update_begin(int sample_every, int max_samples);
int ticc_profile_update_end(double* total_ms, double* union_ms, int* count, double* total_bytes, double* total_flop);

#ifdef __cplusplus
}
#endif

#endif /* TICC_HIP_H */
