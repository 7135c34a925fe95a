/*
 * Forward (ancestral) sampling of a discrete Bayesian network into uint8 columns.
 *
 * Host-side data generator for the synthetic workloads.  Every draw comes from a
 * counter-based generator u(seed, variable, row) = splitmix64 finalizer of a mixed key,
 * so rows are independent and the output does not depend on the thread count.
 *
 * Layout written: cols[v * stride + r] = state of variable v in row r (identity coding).
 */
#include <stdint.h>

static inline uint64_t fpgm_mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline double fpgm_unit(uint64_t seed, int64_t v, int64_t r)
{
    const uint64_t k = fpgm_mix64(seed ^ fpgm_mix64(((uint64_t)v << 40) ^ (uint64_t)r));
    return (double)(k >> 11) * (1.0 / 9007199254740992.0);
}

/*
 * order[n_vars]: a topological order.  par_off[n_vars+1] / par[]: parents of each
 * variable (ascending ids); the parent configuration index is Horner over them in
 * that order.  cpt_off[v]: offset of variable v's cumulative CPT rows in cum[], row
 * length card[v].  The state is the number of cumulative entries (excluding the
 * last) that are <= u.
 */
void fpgm_gen_forward(int32_t n_vars, int64_t n_rows, int64_t stride, const int32_t *order, const int32_t *card,
                      const int64_t *par_off, const int32_t *par, const int64_t *cpt_off, const double *cum,
                      uint64_t seed, uint8_t *cols)
{
    const int64_t B = 4096;
    const int64_t nb = (n_rows + B - 1) / B;
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nb; ++blk) {
        const int64_t r0 = blk * B;
        const int64_t r1 = (r0 + B < n_rows) ? r0 + B : n_rows;
        for (int32_t oi = 0; oi < n_vars; ++oi) {
            const int32_t v = order[oi];
            const int32_t c = card[v];
            const double *cv = cum + cpt_off[v];
            uint8_t *out = cols + (int64_t)v * stride;
            for (int64_t r = r0; r < r1; ++r) {
                int64_t cfg = 0;
                for (int64_t k = par_off[v]; k < par_off[v + 1]; ++k) {
                    const int32_t p = par[k];
                    cfg = cfg * card[p] + cols[(int64_t)p * stride + r];
                }
                const double u = fpgm_unit(seed, v, r);
                const double *row = cv + cfg * c;
                int32_t s = 0;
                for (int32_t k = 0; k < c - 1; ++k) s += (u >= row[k]);
                out[r] = (uint8_t)s;
            }
        }
    }
}

/* Bootstrap resample: out[v, r] = base[v, idx[r]]. */
void fpgm_gen_gather(int32_t n_vars, int64_t n_rows, const uint8_t *base, int64_t base_stride, const int64_t *idx,
                     uint8_t *out, int64_t out_stride)
{
#pragma omp parallel for schedule(static)
    for (int32_t v = 0; v < n_vars; ++v) {
        const uint8_t *b = base + (int64_t)v * base_stride;
        uint8_t *o = out + (int64_t)v * out_stride;
        for (int64_t r = 0; r < n_rows; ++r) o[r] = b[idx[r]];
    }
}
