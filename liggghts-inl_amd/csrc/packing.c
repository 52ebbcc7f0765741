/*
 * packing.c -- seeded dense random sphere packing for synthetic workloads.
 *
 * Host-only helper behind workloads.packed_column: spheres with radii
 * r*shrink are placed uniformly at random in a periodic (Lx, Ly, Lz) box and
 * grown from 0.55x to full size while pairwise overlaps are removed (each
 * overlapping pair is pushed apart along its centre line by half the overlap,
 * Jacobi style).  Deterministic for given inputs.  Not part of the timestep
 * path; it produces the initial state a `read_data` file would provide.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double wrapd(double x, double L)
{
    x = fmod(x, L);
    return x < 0 ? x + L : x;
}

/* pos: n*3 in/out (initial positions in [0,box)); returns iterations used */
int dem_relax_packing(int n, double *pos, const double *r, const double box[3], double shrink,
                      int iters, int grow_steps, double tol)
{
    double rmax = 0.0;
    for (int i = 0; i < n; i++) if (r[i] > rmax) rmax = r[i];
    const double cut = 2.0 * rmax * shrink + 0.15 * rmax;
    int nc[3];
    for (int d = 0; d < 3; d++) {
        nc[d] = (int)(box[d] / cut);
        if (nc[d] < 1) nc[d] = 1;
    }
    const int ncell = nc[0] * nc[1] * nc[2];
    int *head = malloc(sizeof(int) * ncell);
    int *next = malloc(sizeof(int) * n);
    int cap = n * 16, np = 0;
    int *pi = malloc(sizeof(int) * cap), *pj = malloc(sizeof(int) * cap);
    double *disp = malloc(sizeof(double) * 3 * n);
    int it;
    for (it = 0; it < iters; it++) {
        double s = 0.55 + 0.45 * it / (double)grow_steps;
        if (s > 1.0) s = 1.0;
        s *= shrink;
        if (it % 4 == 0) {
            for (int c = 0; c < ncell; c++) head[c] = -1;
            for (int i = 0; i < n; i++) {
                int c3[3];
                for (int d = 0; d < 3; d++) {
                    pos[3 * i + d] = wrapd(pos[3 * i + d], box[d]);
                    int c = (int)(pos[3 * i + d] / box[d] * nc[d]);
                    c3[d] = c >= nc[d] ? nc[d] - 1 : c;
                }
                const int c = (c3[2] * nc[1] + c3[1]) * nc[0] + c3[0];
                next[i] = head[c];
                head[c] = i;
            }
            np = 0;
            for (int i = 0; i < n; i++) {
                int c3[3];
                for (int d = 0; d < 3; d++) {
                    int c = (int)(pos[3 * i + d] / box[d] * nc[d]);
                    c3[d] = c >= nc[d] ? nc[d] - 1 : c;
                }
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            int cc[3] = {c3[0] + dx, c3[1] + dy, c3[2] + dz};
                            int dup = 0;
                            for (int d = 0; d < 3; d++) {
                                if (nc[d] < 3 && (cc[d] < 0 || cc[d] >= nc[d])) dup = 1;
                                cc[d] = (cc[d] + nc[d]) % nc[d];
                            }
                            if (dup) continue;
                            for (int j = head[(cc[2] * nc[1] + cc[1]) * nc[0] + cc[0]]; j >= 0; j = next[j]) {
                                if (j <= i) continue;
                                double dd2 = 0.0;
                                for (int q = 0; q < 3; q++) {
                                    double t = pos[3 * j + q] - pos[3 * i + q];
                                    t -= box[q] * nearbyint(t / box[q]);
                                    dd2 += t * t;
                                }
                                if (dd2 > cut * cut) continue;
                                if (np == cap) {
                                    cap *= 2;
                                    pi = realloc(pi, sizeof(int) * cap);
                                    pj = realloc(pj, sizeof(int) * cap);
                                }
                                pi[np] = i;
                                pj[np] = j;
                                np++;
                            }
                        }
            }
        }
        memset(disp, 0, sizeof(double) * 3 * n);
        double maxov = 0.0;
        for (int k = 0; k < np; k++) {
            const int i = pi[k], j = pj[k];
            double d[3], dist2 = 0.0;
            for (int q = 0; q < 3; q++) {
                double t = pos[3 * j + q] - pos[3 * i + q];
                t -= box[q] * nearbyint(t / box[q]);
                d[q] = t;
                dist2 += t * t;
            }
            const double dist = sqrt(dist2);
            const double ov = (r[i] + r[j]) * s - dist;
            if (ov <= 0.0 || dist == 0.0) continue;
            if (ov > maxov) maxov = ov;
            const double f = 0.5 * 1.05 * ov / dist;
            for (int q = 0; q < 3; q++) {
                disp[3 * i + q] -= f * d[q];
                disp[3 * j + q] += f * d[q];
            }
        }
        if (it >= grow_steps && maxov < tol * rmax) break;
        for (int q = 0; q < 3 * n; q++) pos[q] += disp[q];
    }
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) pos[3 * i + d] = wrapd(pos[3 * i + d], box[d]);
    free(head); free(next); free(pi); free(pj); free(disp);
    return it;
}
