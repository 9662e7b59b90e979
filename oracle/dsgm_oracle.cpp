/*
 * FP64 CPU reference of the differentiable SGM -- TEST INFRASTRUCTURE ONLY.
 * Only tests/ and the benchmark's CPU-baseline leg load it.
 *
 * The differentiable form has no counterpart in the reference (it is a new
 * form beside the plain path).  Definition, per image of a batch:
 *   cost C[H][W][D] (float), penalties P1[H][W], P2[H][W] (per pixel, taken
 *   at the pixel being updated), temperature T, ground truth gt[H][W].
 *   For each direction r (top->bottom, left->right, bottom->top,
 *   right->left) over the whole image:
 *     L(p0, d) = C(p0, d)
 *     L(p, d)  = C(p, d) + min(L'(d), min(L'(d-1), L'(d+1)) + P1(p),
 *                              m' + P2(p)) - m'
 *     with L'(-1) = L'(D) = +inf, m' = min_d L'(d)
 *   S = L_UD + L_LR + L_DU + L_RL (in that order)
 *   prob(p, d) = softmax_d(-S(p, d) / T);  disp(p) = sum_d d prob(p, d)
 *   loss = sum_p |disp(p) - gt(p)| / (H W)
 * Subgradient convention of the min: the first candidate attaining it in the
 * order [L'(d), L'(d-1)+P1, L'(d+1)+P1, m'+P2]; m' = L'(k*) with k* the
 * smallest argmin.  sign(0) = 0.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

inline size_t at3(int x, int y, int d, int W, int D) { return ((size_t)y * W + x) * D + d; }

struct Line {
    std::vector<size_t> pix;   // pixel index (y*W + x) per step
};

std::vector<Line> lines_of(int dir, int W, int H) {
    std::vector<Line> out;
    if (dir == 1 || dir == 3) {
        for (int y = 0; y < H; ++y) {
            Line l;
            for (int i = 0; i < W; ++i) {
                const int x = (dir == 1) ? i : W - 1 - i;
                l.pix.push_back((size_t)y * W + x);
            }
            out.push_back(std::move(l));
        }
    } else {
        for (int x = 0; x < W; ++x) {
            Line l;
            for (int i = 0; i < H; ++i) {
                const int y = (dir == 0) ? i : H - 1 - i;
                l.pix.push_back((size_t)y * W + x);
            }
            out.push_back(std::move(l));
        }
    }
    return out;
}

// forward of one line; keeps L of every step plus winners (0..3) and k*
void forward_line(const double* C, const double* P1, const double* P2, int D, const Line& ln,
                  std::vector<double>& L, std::vector<uint8_t>& win, std::vector<int>& kstar) {
    const size_t n = ln.pix.size();
    L.assign(n * D, 0.0);
    win.assign(n * D, 0);
    kstar.assign(n, 0);
    const double inf = std::numeric_limits<double>::infinity();
    for (int d = 0; d < D; ++d) L[d] = C[ln.pix[0] * D + d];
    for (size_t s = 0; s < n; ++s) {
        const double* cur = &L[s * D];
        int k = 0;
        for (int d = 1; d < D; ++d)
            if (cur[d] < cur[k]) k = d;
        kstar[s] = k;
        if (s + 1 == n) break;
        const double m = cur[k];
        const size_t p = ln.pix[s + 1];
        double* nxt = &L[(s + 1) * D];
        for (int d = 0; d < D; ++d) {
            const double c0 = cur[d];
            const double c1 = (d > 0 ? cur[d - 1] : inf) + P1[p];
            const double c2 = (d + 1 < D ? cur[d + 1] : inf) + P1[p];
            const double c3 = m + P2[p];
            double best = c0;
            int w = 0;
            if (c1 < best) { best = c1; w = 1; }
            if (c2 < best) { best = c2; w = 2; }
            if (c3 < best) { best = c3; w = 3; }
            win[(s + 1) * D + d] = (uint8_t)w;
            nxt[d] = C[p * D + d] + best - m;
        }
    }
}

}  // namespace

extern "C" {

/* Forward: S[H][W][D] (sum of the four directions), disp[H][W]; returns loss
 * (gt may be NULL: loss 0). */
double dsgm_forward(const double* C, const double* P1, const double* P2, const double* gt,
                    int H, int W, int D, double T, double* S, double* disp) {
    const size_t npix = (size_t)H * W;
    std::memset(S, 0, npix * D * sizeof(double));
    for (int r = 0; r < 4; ++r) {
        std::vector<Line> lines = lines_of(r, W, H);
#pragma omp parallel for schedule(dynamic)
        for (int li = 0; li < (int)lines.size(); ++li) {
            std::vector<double> L;
            std::vector<uint8_t> win;
            std::vector<int> ks;
            forward_line(C, P1, P2, D, lines[li], L, win, ks);
            for (size_t s = 0; s < lines[li].pix.size(); ++s)
                for (int d = 0; d < D; ++d) S[lines[li].pix[s] * D + d] += L[s * D + d];
        }
    }
    double loss = 0.0;
#pragma omp parallel for reduction(+ : loss)
    for (long long p = 0; p < (long long)npix; ++p) {
        const double* s = S + p * D;
        double mn = s[0];
        for (int d = 1; d < D; ++d) mn = std::min(mn, s[d]);
        double z = 0.0, acc = 0.0;
        for (int d = 0; d < D; ++d) {
            const double e = std::exp(-(s[d] - mn) / T);
            z += e;
            acc += d * e;
        }
        disp[p] = acc / z;
        if (gt) loss += std::fabs(disp[p] - gt[p]);
    }
    return loss / (double)npix;
}

/* Backward of the L1 loss: gC[H][W][D], gP1[H][W], gP2[H][W]; returns loss. */
double dsgm_backward(const double* C, const double* P1, const double* P2, const double* gt,
                     int H, int W, int D, double T, double* gC, double* gP1, double* gP2) {
    const size_t npix = (size_t)H * W;
    std::vector<double> S(npix * D), disp(npix), gS(npix * D);
    const double loss = dsgm_forward(C, P1, P2, gt, H, W, D, T, S.data(), disp.data());
#pragma omp parallel for
    for (long long p = 0; p < (long long)npix; ++p) {
        const double* s = &S[p * D];
        double mn = s[0];
        for (int d = 1; d < D; ++d) mn = std::min(mn, s[d]);
        double z = 0.0;
        for (int d = 0; d < D; ++d) z += std::exp(-(s[d] - mn) / T);
        const double diff = disp[p] - gt[p];
        const double gd = (diff > 0 ? 1.0 : (diff < 0 ? -1.0 : 0.0)) / (double)npix;
        for (int d = 0; d < D; ++d) {
            const double pr = std::exp(-(s[d] - mn) / T) / z;
            gS[p * D + d] = gd * (-1.0 / T) * pr * (d - disp[p]);
        }
    }
    std::memset(gC, 0, npix * D * sizeof(double));
    std::memset(gP1, 0, npix * sizeof(double));
    std::memset(gP2, 0, npix * sizeof(double));
    for (int r = 0; r < 4; ++r) {
        std::vector<Line> lines = lines_of(r, W, H);
#pragma omp parallel for schedule(dynamic)
        for (int li = 0; li < (int)lines.size(); ++li) {
            const Line& ln = lines[li];
            std::vector<double> L;
            std::vector<uint8_t> win;
            std::vector<int> ks;
            forward_line(C, P1, P2, D, ln, L, win, ks);
            const size_t n = ln.pix.size();
            std::vector<double> G(D), Gp(D);
            for (int d = 0; d < D; ++d) G[d] = gS[ln.pix[n - 1] * D + d];
            for (size_t s = n - 1; s > 0; --s) {
                const size_t p = ln.pix[s];
                double sum_all = 0.0, sum3 = 0.0, sum12 = 0.0;
                for (int d = 0; d < D; ++d) {
                    gC[p * D + d] += G[d];
                    const int w = win[s * D + d];
                    sum_all += G[d];
                    if (w == 3) sum3 += G[d];
                    if (w == 1 || w == 2) sum12 += G[d];
                }
                gP1[p] += sum12;
                gP2[p] += sum3;
                const size_t pp = ln.pix[s - 1];
                for (int d = 0; d < D; ++d) {
                    double g = gS[pp * D + d];
                    if (win[s * D + d] == 0) g += G[d];
                    if (d + 1 < D && win[s * D + d + 1] == 1) g += G[d + 1];
                    if (d > 0 && win[s * D + d - 1] == 2) g += G[d - 1];
                    if (d == ks[s - 1]) g += sum3 - sum_all;
                    Gp[d] = g;
                }
                std::swap(G, Gp);
            }
            for (int d = 0; d < D; ++d) gC[ln.pix[0] * D + d] += G[d];
        }
    }
    return loss;
}

}  // extern "C"
