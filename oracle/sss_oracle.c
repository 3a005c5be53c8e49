/*
 * ORACLE — test infrastructure only.  CPU float64 restatement of the
 * reference's TOF single-scatter-simulation path, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * Nothing in the shipped library calls, links or loads this file.
 *
 * Followed, statement by statement, with the float64 operation order kept:
 *   ora_trace        Python/openSSS/RayTracing3DTOF.py:29-180 (+ BoxIntersectTest
 *                    :204-269, sub2ind :182-202)
 *   ora_sss          Python/openSSS/SingleScatterSimulationTOF.py:300-559 (bodySSS)
 *                    and the slice sum of :193-201
 *   ora_line_integrals  Python/openSSS/TailsMask.py:84-163 trace + sums
 *   ora_backproject  Python/openSSS/BackProjection.py:31-113 / :217-300
 *
 * Build with -ffp-contract=off (no fused multiply-add), as numba and numpy do.
 *
 * Parity pins (tests/): exact-rational brute-force tracer on small grids,
 * analytic segment/box length, the s_k recurrence, tie and axis-parallel
 * cases; closed-form single-voxel SSS phantom, TOF partition of unity,
 * activity linearity, mirror symmetry.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* --------------------------------------------------------------------- */
/* Ray tracer                                                             */
/* --------------------------------------------------------------------- */

static int box_intersect(const double* gb, const double* lc, double* tmin_out, double* tmax_out)
{
    double x_div = 1 / (lc[3] - lc[0] + 1e-20);
    double t_min, t_max;
    if (x_div >= 0) { t_min = (gb[0] - lc[0]) * x_div; t_max = (gb[3] - lc[0]) * x_div; }
    else            { t_min = (gb[3] - lc[0]) * x_div; t_max = (gb[0] - lc[0]) * x_div; }
    double y_div = 1 / (lc[4] - lc[1] + 1e-20);
    double y_t_min, y_t_max;
    if (y_div >= 0) { y_t_min = (gb[1] - lc[1]) * y_div; y_t_max = (gb[4] - lc[1]) * y_div; }
    else            { y_t_min = (gb[4] - lc[1]) * y_div; y_t_max = (gb[1] - lc[1]) * y_div; }
    if (t_min > y_t_max || y_t_min > t_max) return 0;
    if (y_t_min > t_min) t_min = y_t_min;
    if (y_t_max < t_max) t_max = y_t_max;
    double z_div = 1 / (lc[5] - lc[2] + 1e-20);
    double z_t_min, z_t_max;
    if (z_div >= 0) { z_t_min = (gb[2] - lc[2]) * z_div; z_t_max = (gb[5] - lc[2]) * z_div; }
    else            { z_t_min = (gb[5] - lc[2]) * z_div; z_t_max = (gb[2] - lc[2]) * z_div; }
    if (t_min > z_t_max || z_t_min > t_max) return 0;
    if (z_t_min > t_min) t_min = z_t_min;
    if (z_t_max < t_max) t_max = z_t_max;
    if ((t_min >= 1 && t_max >= 1) || (t_min <= 0 && t_max <= 0)) return 0;
    *tmin_out = t_min;
    *tmax_out = t_max;
    return 1;
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Returns the voxel count n (0 if the segment misses the box, -1 if the
 * walk would exceed cap).  idx is the reference's 1-based index
 * X + (Y-1)*Ny + (Z-1)*Ny*Nx. */
int ora_trace(const int32_t* gs, const double* gb, const double* lc,
              double* L, int32_t* idx, double* s, int cap)
{
    double tMin, tMax;
    if (!box_intersect(gb, lc, &tMin, &tMax)) return 0;
    tMin = tMin > 0 ? tMin : 0;   /* max(tMin, 0) */
    tMax = tMax < 1 ? tMax : 1;   /* min(tMax, 1) */
    const double Nx = gs[0], Ny = gs[1], Nz = gs[2];
    double xv = (gb[3] - gb[0]) / gs[0];
    double yv = (gb[4] - gb[1]) / gs[1];
    double zv = (gb[5] - gb[2]) / gs[2];
    double xVec = lc[3] - lc[0], yVec = lc[4] - lc[1], zVec = lc[5] - lc[2];
    double xs = lc[0] + xVec * tMin, ys = lc[1] + yVec * tMin, zs = lc[2] + zVec * tMin;
    double xe = lc[0] + xVec * tMax, ye = lc[1] + yVec * tMax, ze = lc[2] + zVec * tMax;

    double X = fmax(1, fmin(Nx, ceil((xs - gb[0]) / xv)));
    double xEnd = fmax(1, fmin(Nx, ceil((xe - gb[0]) / xv)));
    double xStep, xTD, xTM;
    if (xVec > 0) { xStep = 1; xTD = xv / xVec; xTM = tMin + (gb[0] + X * xv - xs) / xVec; }
    else if (xVec < 0) { xStep = -1; xTD = xv / -xVec; xTM = tMin + (gb[0] + (X - 1) * xv - xs) / xVec; }
    else { xStep = 0; xTM = tMax; xTD = tMax; }

    double Y = fmax(1, fmin(Ny, ceil((ys - gb[1]) / yv)));
    double yEnd = fmax(1, fmin(Ny, ceil((ye - gb[1]) / yv)));
    double yStep, yTD, yTM;
    if (yVec > 0) { yStep = 1; yTD = yv / yVec; yTM = tMin + (gb[1] + Y * yv - ys) / yVec; }
    else if (yVec < 0) { yStep = -1; yTD = yv / -yVec; yTM = tMin + (gb[1] + (Y - 1) * yv - ys) / yVec; }
    else { yStep = 0; yTM = tMax; yTD = tMax; }

    double Z = clampd(ceil((zs - gb[2]) / zv), 1, Nz);
    double zEnd = clampd(ceil((ze - gb[2]) / zv), 1, Nz);
    double zStep, zTD, zTM;
    if (zVec > 0) { zStep = 1; zTD = zv / zVec; zTM = tMin + (gb[2] + Z * zv - zs) / zVec; }
    else if (zVec < 0) { zStep = -1; zTD = zv / -zVec; zTM = tMin + (gb[2] + (Z - 1) * zv - zs) / zVec; }
    else { zStep = 0; zTM = tMax; zTD = tMax; }

    /* A float64 end voxel can disagree with where the steps lead when the
     * segment runs (almost) inside a voxel face; the reference walk then
     * never reaches its end voxel (no termination, writes past its arrays).
     * Both this restatement and the device tracer stop after
     * |dX|+|dY|+|dZ|+1 voxels -- the count every terminating walk has -- and
     * close the last voxel at tMax. */
    const int n_max = (int)(fabs(xEnd - X) + fabs(yEnd - Y) + fabs(zEnd - Z)) + 1;
    if (n_max > cap) return -1;
    idx[0] = (int32_t)(X + (Y - 1) * Ny + (Z - 1) * Ny * Nx);
    int n = 1;
    double px = xVec * tMin, py = yVec * tMin, pz = zVec * tMin;
    double prev = 0;
    while ((X != xEnd || Y != yEnd || Z != zEnd) && n < n_max) {
        double len;
        /* same axis choice as the reference; a step that would leave the grid
         * (only possible in the non-terminating case above) ends the walk */
        int axis = (xTM < yTM) ? (xTM < zTM ? 0 : 2) : (yTM < zTM ? 1 : 2);
        double nX = axis == 0 ? X + xStep : X, nY = axis == 1 ? Y + yStep : Y, nZ = axis == 2 ? Z + zStep : Z;
        if (nX < 1 || nX > Nx || nY < 1 || nY > Ny || nZ < 1 || nZ > Nz) break;
        if (axis == 0) { X = nX; len = xTM; xTM += xTD; }
        else if (axis == 1) { Y = nY; len = yTM; yTM += yTD; }
        else { Z = nZ; len = zTM; zTM += zTD; }
        double dx = px - xVec * len, dy = py - yVec * len, dz = pz - zVec * len;
        L[n - 1] = sqrt(dx * dx + dy * dy + dz * dz);
        s[n - 1] = prev + L[n - 1] / 2;
        prev = s[n - 1];
        px = xVec * len; py = yVec * len; pz = zVec * len;
        n += 1;
        idx[n - 1] = (int32_t)(X + (Y - 1) * Ny + (Z - 1) * Ny * Nx);
    }
    double dx = px - xVec * tMax, dy = py - yVec * tMax, dz = pz - zVec * tMax;
    L[n - 1] = sqrt(dx * dx + dy * dy + dz * dz);
    s[n - 1] = prev + L[n - 1] / 2;
    return n;
}

/* --------------------------------------------------------------------- */
/* Single scatter simulation                                              */
/* --------------------------------------------------------------------- */

typedef struct {
    int32_t grid[3];          /* X, Y, Z */
    double bounds[6];         /* ImageSize, cm */
    const double* mu;         /* F-order flat, cm^-1 */
    const double* lam;        /* F-order flat */
    int32_t n_rings, n_det;   /* full scanner */
    const double* geom;       /* (NR, N, 3) C-order, cm */
    const double* normals;    /* (NR, N, 3) */
    double det_size[2];       /* mm */
    const double* ratio;      /* (n_energy,) water ratio */
    const double* eff;        /* (n_energy,) efficiency */
    int32_t n_energy;
    int32_t ref_row;          /* 511 keV row */
    const double* tof;        /* (B, 8001) */
    int32_t n_bins;
    int32_t nru, ndu;
    const int32_t* rings;     /* (NRU,) ring ids */
    const int32_t* dets;      /* (NDU,) detector ids */
    const int16_t* sino;      /* (N, N, 2) (angular, radial) */
    const int32_t* cell_of;   /* ((N+1)*(N/2),) compact cell id or -1 */
    int32_t n_cells;          /* compact cells per plane */
    int32_t n_sp;
    const double* sp_xyz;     /* (n_sp, 3) scatter-point positions, slice-sorted */
    int32_t n_slices;
    const int32_t* slice_ptr; /* (n_slices+1,) */
    int32_t n_threads;
    int64_t* stat_offsets_out_of_range;  /* optional counter */
} ora_sss_args;

#define ORA_TOF_HALF 4000
#define ORA_LIGHT_CM_PS (299792458.0 * 1e-10)
#define ORA_E_REF_J (511.0E3 * 1.6E-19)
#define ORA_ME 9.10938356E-31
#define ORA_C_M_S 299792458.0

static double ora_kn(double energy, double angle_deg)
{
    double gamma = energy / (ORA_ME * (ORA_C_M_S * ORA_C_M_S));
    double ratio = 1.0 / (1.0 + gamma * (1.0 - cos(angle_deg * (M_PI / 180.0))));
    double sn = sin(angle_deg * (M_PI / 180.0));
    return (ratio * ratio) * (ratio + 1 / ratio - sn * sn);
}

typedef struct {
    int n;
    double* act_ray;   /* lambda * L */
    double* s;
} ora_half;

/* One normalisation slice: accumulate into sum/cnt (f32, as the reference). */
static void ora_slice(const ora_sss_args* a, int slice, float* sum, float* cnt,
                      double* tbuf_L, int32_t* tbuf_i, double* tbuf_s,
                      ora_half* hp, double* linepaths, double* angles, int64_t* oor)
{
    const int X = a->grid[0], Y = a->grid[1], Z = a->grid[2];
    const int cap = X + Y + Z;
    const int NRU = a->nru, NDU = a->ndu, N = a->n_det, B = a->n_bins;
    const int NA = N / 2;
    const size_t plane_stride = (size_t)a->n_cells;
    const size_t bin_stride = (size_t)NRU * NRU * plane_stride;
    const double light = ORA_LIGHT_CM_PS;
    double gs[3] = {X, Y, Z};
    (void)gs;
    for (int sp = a->slice_ptr[slice]; sp < a->slice_ptr[slice + 1]; ++sp) {
        const double xS = a->sp_xyz[3 * sp], yS = a->sp_xyz[3 * sp + 1], zS = a->sp_xyz[3 * sp + 2];
        /* half paths S -> detectors (SingleScatterSimulationTOF.py:400-430) */
        for (int r = 0; r < NRU; ++r) {
            const int ring = a->rings[r];
            const double zD = a->geom[((size_t)ring * N + 0) * 3 + 2];
            for (int i = 0; i < NDU; ++i) {
                const int det = a->dets[i];
                const double xD = a->geom[((size_t)ring * N + det) * 3 + 0];
                const double yD = a->geom[((size_t)ring * N + det) * 3 + 1];
                double lc[6] = {xS, yS, zS, xD, yD, zD};
                int n = ora_trace(a->grid, a->bounds, lc, tbuf_L, tbuf_i, tbuf_s, cap);
                if (n < 0) n = 0;
                ora_half* h = &hp[r * NDU + i];
                h->n = n;
                double act = 0, att = 0;
                for (int k = 0; k < n; ++k) {
                    int id = tbuf_i[k] - 1;
                    double ar = a->lam[id] * tbuf_L[k];
                    h->act_ray[k] = ar;
                    h->s[k] = tbuf_s[k];
                    act += ar;
                    att += a->mu[id] * tbuf_L[k];
                }
                linepaths[(r * NDU + i) * 2 + 0] = exp(-att);
                linepaths[(r * NDU + i) * 2 + 1] = act;
                const double* nv = &a->normals[((size_t)ring * N + det) * 3];
                double sv0 = xD - xS, sv1 = yD - yS, sv2 = zD - zS;
                double dot = sv0 * nv[0] + sv1 * nv[1] + sv2 * nv[2];
                double nn = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
                double ns = sqrt(sv0 * sv0 + sv1 * sv1 + sv2 * sv2);
                angles[r * NDU + i] = fabs(acos(dot / (nn * ns)) * (180.0 / M_PI));
            }
        }
        /* LOR double loop (:433-553) */
        for (int r1 = 0; r1 < NRU; ++r1) {
            const int ring1 = a->rings[r1];
            const double z1 = a->geom[((size_t)ring1 * N + 0) * 3 + 2];
            for (int i1 = 0; i1 < NDU; ++i1) {
                const int D1 = a->dets[i1];
                const double x1 = a->geom[((size_t)ring1 * N + D1) * 3 + 0];
                const double y1 = a->geom[((size_t)ring1 * N + D1) * 3 + 1];
                const double sv1 = sqrt((x1 - xS) * (x1 - xS) + (y1 - yS) * (y1 - yS) + (z1 - zS) * (z1 - zS));
                const double A1 = linepaths[(r1 * NDU + i1) * 2 + 0];
                const double act1 = linepaths[(r1 * NDU + i1) * 2 + 1];
                if (A1 == 0) continue;
                const ora_half* h1 = &hp[r1 * NDU + i1];
                for (int r2 = 0; r2 < NRU; ++r2) {
                    const int ring2 = a->rings[r2];
                    const double z2 = a->geom[((size_t)ring2 * N + 0) * 3 + 2];
                    for (int i2 = 0; i2 < NDU; ++i2) {
                        const int D2 = a->dets[i2];
                        if ((D1 == D2 && r2 == r1) || D1 > D2) continue;
                        const double x2 = a->geom[((size_t)ring2 * N + D2) * 3 + 0];
                        const double y2 = a->geom[((size_t)ring2 * N + D2) * 3 + 1];
                        const double sv2 = sqrt((x2 - xS) * (x2 - xS) + (y2 - yS) * (y2 - yS) + (z2 - zS) * (z2 - zS));
                        const double lor = ((x1 - x2) * (x1 - x2)) + ((y1 - y2) * (y1 - y2)) + ((z1 - z2) * (z1 - z2));
                        const double A2 = linepaths[(r2 * NDU + i2) * 2 + 0];
                        const double act2 = linepaths[(r2 * NDU + i2) * 2 + 1];
                        float counts;
                        if (A2 == 0 || (act1 == 0 && act2 == 0)) continue;
                        else if (act1 != 0 && act2 != 0) counts = 2;
                        else counts = 1;
                        const int ang = a->sino[((size_t)D1 * N + D2) * 2 + 0];
                        const int rad = a->sino[((size_t)D1 * N + D2) * 2 + 1];
                        /* CalcAngleScatter (:215-235) */
                        const double s1q = sv1 * sv1, s2q = sv2 * sv2;
                        const double theta = 180 - acos((s1q + s2q - lor) / (2 * sqrt(s1q * s2q))) * (180.0 / M_PI);
                        const double e_sc = floor(511.0 / (1 + (511.0 / 511.0) * (1 - cos(theta * (M_PI / 180.0)))) + 0.5);
                        const int k = (int)(e_sc * 2) - 1;
                        const double eff = a->eff[k] * a->eff[a->ref_row];
                        if (eff == 0) continue;
                        const double rr = a->ratio[k];
                        const double A1s = pow(A1, rr), A2s = pow(A2, rr);
                        const double dist = (sv1 - sv2) / 2;
                        double ab1[64], ab2[64];
                        for (int b = 0; b < B; ++b) { ab1[b] = 0; ab2[b] = 0; }
                        if (act1 != 0) {
                            for (int kk = 0; kk < h1->n; ++kk) {
                                double off = floor(2 * (dist - h1->s[kk]) / light + 0.5) + ORA_TOF_HALF;
                                long o = (long)off;
                                if (o < 0 || o > 2 * ORA_TOF_HALF) { if (oor) (*oor)++; continue; }
                                for (int b = 0; b < B; ++b) ab1[b] += h1->act_ray[kk] * a->tof[(size_t)b * (2 * ORA_TOF_HALF + 1) + o];
                            }
                        }
                        if (act2 != 0) {
                            const ora_half* h2 = &hp[r2 * NDU + i2];
                            for (int kk = 0; kk < h2->n; ++kk) {
                                double off = floor(2 * (dist + h2->s[kk]) / light + 0.5) + ORA_TOF_HALF;
                                long o = (long)off;
                                if (o < 0 || o > 2 * ORA_TOF_HALF) { if (oor) (*oor)++; continue; }
                                for (int b = 0; b < B; ++b) ab2[b] += h2->act_ray[kk] * a->tof[(size_t)b * (2 * ORA_TOF_HALF + 1) + o];
                            }
                        }
                        const double c1 = fabs(cos(angles[r1 * NDU + i1] * (M_PI / 180.0)));
                        const double c2 = fabs(cos(angles[r2 * NDU + i2] * (M_PI / 180.0)));
                        const double ds = a->det_size[0] * a->det_size[1] * 1e-2;
                        const double geff = ((ds * ds) * c1) * (c2 / (4 * M_PI * s1q * s2q));
                        const double pre = geff * eff * ora_kn(ORA_E_REF_J, theta);
                        const int plane = r2 + r1 * NRU;
                        const int cell = a->cell_of[(size_t)rad * NA + ang];
                        float* dst = sum + (size_t)plane * plane_stride + cell;
                        for (int b = 0; b < B; ++b) {
                            double p = pre * (A1 * A2s * ab1[b] + A2 * A1s * ab2[b]);
                            dst[b * bin_stride] = (float)((double)dst[b * bin_stride] + p);
                        }
                        cnt[(size_t)plane * plane_stride + cell] += counts;
                    }
                }
            }
        }
    }
}

/* Output: out[b][plane][cell] float32 = sum over slices of nan_to_num(Sum/Cnt). */
int ora_sss(const ora_sss_args* a, float* out)
{
    const int NRU = a->nru, NDU = a->ndu, B = a->n_bins;
    if (B > 64) return -2;
    if (a->grid[0] != a->grid[1]) return -3;
    const size_t cells = (size_t)NRU * NRU * a->n_cells;
    const size_t total = (size_t)B * cells;
    memset(out, 0, total * sizeof(float));
    int nt = a->n_threads > 0 ? a->n_threads : 1;
#ifdef _OPENMP
    omp_set_num_threads(nt);
#else
    nt = 1;
#endif
    const int cap = a->grid[0] + a->grid[1] + a->grid[2];
    float* sums = (float*)malloc((size_t)nt * total * sizeof(float));
    float* cnts = (float*)malloc((size_t)nt * cells * sizeof(float));
    if (!sums || !cnts) { free(sums); free(cnts); return -1; }
    int64_t oor_total = 0;
    for (int base = 0; base < a->n_slices; base += nt) {
        int nb = a->n_slices - base < nt ? a->n_slices - base : nt;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : oor_total)
        for (int j = 0; j < nb; ++j) {
            float* sum = sums + (size_t)j * total;
            float* cnt = cnts + (size_t)j * cells;
            memset(sum, 0, total * sizeof(float));
            memset(cnt, 0, cells * sizeof(float));
            double* tL = (double*)malloc(cap * sizeof(double));
            int32_t* tI = (int32_t*)malloc(cap * sizeof(int32_t));
            double* tS = (double*)malloc(cap * sizeof(double));
            ora_half* hp = (ora_half*)malloc((size_t)NRU * NDU * sizeof(ora_half));
            double* pool = (double*)malloc((size_t)NRU * NDU * cap * 2 * sizeof(double));
            for (int h = 0; h < NRU * NDU; ++h) {
                hp[h].act_ray = pool + (size_t)h * cap * 2;
                hp[h].s = pool + (size_t)h * cap * 2 + cap;
            }
            double* lp = (double*)malloc((size_t)NRU * NDU * 2 * sizeof(double));
            double* ang = (double*)malloc((size_t)NRU * NDU * sizeof(double));
            int64_t oor = 0;
            ora_slice(a, base + j, sum, cnt, tL, tI, tS, hp, lp, ang, &oor);
            oor_total += oor;
            /* ScatterSlice / ScatterCounts, nan -> 0 (:555-557) */
            for (int b = 0; b < B; ++b)
                for (size_t c = 0; c < cells; ++c) {
                    float q = sum[(size_t)b * cells + c] / cnt[c];
                    sum[(size_t)b * cells + c] = isnan(q) ? 0.0f : q;
                }
            free(tL); free(tI); free(tS); free(hp); free(pool); free(lp); free(ang);
        }
        for (int j = 0; j < nb; ++j) {   /* slice sum in slice order (:201) */
            const float* sum = sums + (size_t)j * total;
            for (size_t e = 0; e < total; ++e) out[e] += sum[e];
        }
    }
    if (a->stat_offsets_out_of_range) *a->stat_offsets_out_of_range = oor_total;
    free(sums); free(cnts);
    return 0;
}

/* --------------------------------------------------------------------- */
/* Line integrals of detector pairs (tail mask / forward projection)      */
/* --------------------------------------------------------------------- */

/* For each segment (6 doubles): n voxels, sum(mu L), sum(lam L) in the
 * reference's numpy order (np.sum: pairwise; the sums here are sequential,
 * which only matters for the last bit of non-zero sums). */
void ora_line_integrals(const int32_t* gs, const double* gb, const double* mu, const double* lam,
                        const double* segs, int64_t n_segs, int32_t* n_out,
                        double* mu_int, double* lam_int)
{
    const int cap = gs[0] + gs[1] + gs[2];
#pragma omp parallel
    {
        double* tL = (double*)malloc(cap * sizeof(double));
        int32_t* tI = (int32_t*)malloc(cap * sizeof(int32_t));
        double* tS = (double*)malloc(cap * sizeof(double));
#pragma omp for schedule(static)
        for (int64_t q = 0; q < n_segs; ++q) {
            int n = ora_trace(gs, gb, segs + 6 * q, tL, tI, tS, cap);
            if (n < 0) n = 0;
            double m = 0, l = 0;
            for (int k = 0; k < n; ++k) {
                m += mu[tI[k] - 1] * tL[k];
                l += lam[tI[k] - 1] * tL[k];
            }
            n_out[q] = n;
            mu_int[q] = m;
            lam_int[q] = l;
        }
        free(tL); free(tI); free(tS);
    }
}

/* Backprojection of one segment list with per-segment values (float32
 * images, as BackProjection.py:68-111).  mode 0: value / n per voxel
 * (no span); mode 1: value * L / sum(L) (span, :294-297).  Segments are
 * processed in the given order; images are accumulated in float32. */
void ora_backproject(const int32_t* gs, const double* gb, const double* segs, int64_t n_segs,
                     const float* v_prompt, const float* v_scatter, const float* v_random, int mode,
                     float* img_prompt, float* img_scatter, float* img_random, float* img_sens)
{
    const int cap = gs[0] + gs[1] + gs[2];
    double* tL = (double*)malloc(cap * sizeof(double));
    int32_t* tI = (int32_t*)malloc(cap * sizeof(int32_t));
    double* tS = (double*)malloc(cap * sizeof(double));
    for (int64_t q = 0; q < n_segs; ++q) {
        if (v_scatter[q] == 0) continue;
        int n = ora_trace(gs, gb, segs + 6 * q, tL, tI, tS, cap);
        if (n <= 0) continue;
        double lsum = 0;
        if (mode == 1) for (int k = 0; k < n; ++k) lsum += tL[k];
        for (int k = 0; k < n; ++k) {
            int id = tI[k] - 1;
            if (mode == 0) {
                img_prompt[id] = img_prompt[id] + v_prompt[q] / (float)n;
                img_scatter[id] = img_scatter[id] + v_scatter[q] / (float)n;
                img_random[id] = img_random[id] + v_random[q] / (float)n;
            } else {
                img_prompt[id] = (float)((double)img_prompt[id] + (double)v_prompt[q] * tL[k] / lsum);
                img_scatter[id] = (float)((double)img_scatter[id] + (double)v_scatter[q] * tL[k] / lsum);
                img_random[id] = (float)((double)img_random[id] + (double)v_random[q] * tL[k] / lsum);
            }
            img_sens[id] = img_sens[id] + 1.0f;
        }
    }
    free(tL); free(tI); free(tS);
}
