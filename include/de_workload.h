/* Synthetic ensemble workloads and device self-checks.
 *
 * No counterpart in the reference crate (it solves one user-built problem per
 * call).  These functions define, once, the seeded ensembles the tests and the
 * benchmark run, so that the CPU oracle and the GPU integrate identical inputs.
 * All generators run on the host; randomness is Philox4x32-10 keyed by
 * (seed, item, stream) with the 53-bit uniform and Box-Muller normal transform
 * shared with the device noise source.
 */
#ifndef DE_WORKLOAD_H
#define DE_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lorenz ensemble: sigma = 10, beta = 8/3, rho = 50 u, y0 = (1,1,1) + 0.1 z.
 * params[3 * n] (sigma, rho, beta rows), y0[3 * n]. */
void de_gen_lorenz(uint64_t seed, int64_t n, double* params, double* y0);

/* Van der Pol: mu = exp(ln 0.1 + u ln 100), y0 = (2, 0) + 0.1 z. */
void de_gen_vdp(uint64_t seed, int64_t n, double* params, double* y0);

/* Spring-chain sizes: N = floor(exp(ln 4 + u (ln 257 - ln 4))) in [4, 256]. */
void de_gen_spring_sizes(uint64_t seed, int64_t n, int32_t* sizes);

/* Prefix offsets for ragged spring chains: state 2N, params N + 2 per system.
 * state_offset and param_offset hold n + 1 entries. */
void de_spring_offsets(int64_t n, const int32_t* sizes, int64_t* state_offset,
                       int64_t* param_offset);

/* Spring-chain parameters and initial state: c ~ U[0, 0.1],
 * k_j = exp(0.5 z) (LogNormal(0, 0.5)), x_j ~ N(0, 1), v_j = 0. */
void de_gen_spring(uint64_t seed, int64_t n, const int32_t* sizes,
                   const int64_t* state_offset, const int64_t* param_offset,
                   double* params, double* y0);

/* Ornstein-Uhlenbeck: per path and component theta = 0.5 + 1.5 u, mu = z,
 * sigma = 0.1 + 0.9 u; y0 = 0.  params[3 * dim * n], y0[dim * n]. */
void de_gen_ou(uint64_t seed, int64_t n, int32_t dim, double* params, double* y0);

/* Geometric Brownian motion as a linear diagonal SDE of dimension 1:
 * a = mu, b = sigma, y0 = s0 for every path.  params[2 * n], y0[n]. */
void de_gen_gbm(int64_t n, double mu, double sigma, double s0, double* params,
                double* y0);

/* Philox4x32-10 block for counter ctr[4] and key (k0, k1). */
void de_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]);

/* Standard normal of the SDE noise source for (seed, path, step, component). */
double de_noise_normal(uint64_t seed, uint64_t path, uint32_t step, uint32_t comp);

/* Device self-checks (run on the GPU, compare with the host in tests):
 * out[i] = pow(x[i], y[i]) with the integrator's pow; x/y/out host memory. */
int de_diag_pow(const double* x, const double* y, double* out, int64_t n);
/* out[3*i + 0..2] = a/b, sqrt(a), fma(a, b, c) as computed on the device. */
int de_diag_arith(const double* a, const double* b, const double* c, double* out,
                  int64_t n);
/* Rust's `{:?}` text of x[i] by the device CSV formatter: bytes out[32 i ..],
 * length len[i] (negative if its digit search reached a loop cap). */
int de_diag_fmt(const double* x, char* out, int32_t* len, int64_t n);
/* out[i] = de_noise_normal(seed, path[i], step[i], comp[i]) on the device. */
int de_diag_noise(uint64_t seed, const uint64_t* path, const uint32_t* step,
                  const uint32_t* comp, double* out, int64_t n);

/* Spring-chain autotuner decision for the size class whose largest chain has
 * `cap` masses (4, 8, ..., 256): lanes per system G, masses per lane C and
 * block size.  Returns 0, or -1 if the class has not been tuned yet. */
int de_spring_tuned(int32_t cap, int32_t* g, int32_t* c, int32_t* block);

#ifdef __cplusplus
}
#endif

#endif /* DE_WORKLOAD_H */
