/*
 * gnnref.c — FP64 CPU oracle of the learned Poisson initial guess (test infrastructure
 * only; the reference has no counterpart, this restates the network the HIP kernels
 * implement).
 *
 * Network over the cell graph (neighbours = cells coupled by the Laplacian):
 *   h0_i  = relu(W_enc^T x_i + b_enc)                       x_i: F_in features
 *   for each layer l:
 *     m_i   = mean_{j in N(i)} h_j
 *     z_i   = W_l^T [h_i ; m_i] + b_l                        W_l: 2H x H
 *     h_i  <- h_i + relu(z_i)
 *   out_i = w_dec . h_i + b_dec
 *   loss  = (1/N) sum_i (out_i - t_i)^2
 * Parameter vector layout: W_enc[F_in][H], b_enc[H], then per layer W_l[2H][H], b_l[H],
 * then w_dec[H], b_dec.
 *
 * Pinned by central finite differences of the loss (tests/test_oracle_gnn.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

int ref_gnn_n_params(int fin, int H, int L) { return fin * H + H + L * (2 * H * H + H) + H + 1; }

/* nb: [n*K] neighbour cells (-1 unused).  Returns the loss; grad (may be NULL) receives
 * dloss/dtheta; out (may be NULL) receives the network output. */
static double gnn_core(int n, int K, const int* nb, int fin, const double* x, const double* t, int H, int L,
                       const double* th, const double* dout_ext, double* grad, double* out);

double ref_gnn_loss_grad(int n, int K, const int* nb, int fin, const double* x, const double* t, int H, int L,
                         const double* th, double* grad, double* out) {
    return gnn_core(n, K, nb, fin, x, t, H, L, th, NULL, grad, out);
}

/* forward (out) and, when dout is given, the gradient of any loss with dloss/dout = dout */
void ref_gnn_forward_backward(int n, int K, const int* nb, int fin, const double* x, int H, int L,
                              const double* th, const double* dout, double* grad, double* out) {
    double* t = (double*)calloc(n, sizeof(double));
    gnn_core(n, K, nb, fin, x, t, H, L, th, dout, dout ? grad : NULL, out);
    free(t);
}

static double gnn_core(int n, int K, const int* nb, int fin, const double* x, const double* t, int H, int L,
                       const double* th, const double* dout_ext, double* grad, double* out) {
    const double* Wenc = th;
    const double* benc = Wenc + fin * H;
    const double* lay = benc + H;
    const double* wdec = lay + L * (2 * H * H + H);
    const double bdec = wdec[H];
    double* h = (double*)malloc(sizeof(double) * (size_t)(L + 1) * n * H);  /* h_l */
    double* m = (double*)malloc(sizeof(double) * (size_t)L * n * H);        /* m_l */
    double* z = (double*)malloc(sizeof(double) * (size_t)(L + 1) * n * H);  /* pre-activations (z[0]: encoder) */
    int* deg = (int*)calloc(n, sizeof(int));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < K; ++k) if (nb[i * K + k] >= 0) deg[i]++;
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < H; ++c) {
            double s = benc[c];
            for (int f = 0; f < fin; ++f) s += x[i * fin + f] * Wenc[f * H + c];
            z[(size_t)i * H + c] = s;
            h[(size_t)i * H + c] = s > 0 ? s : 0;
        }
    for (int l = 0; l < L; ++l) {
        const double* W = lay + l * (2 * H * H + H);
        const double* b = W + 2 * H * H;
        const double* hl = h + (size_t)l * n * H;
        double* ml = m + (size_t)l * n * H;
        double* hn = h + (size_t)(l + 1) * n * H;
        double* zl = z + (size_t)(l + 1) * n * H;
        for (int i = 0; i < n; ++i) {
            for (int c = 0; c < H; ++c) ml[(size_t)i * H + c] = 0.0;
            for (int k = 0; k < K; ++k) {
                int j = nb[i * K + k];
                if (j < 0) continue;
                for (int c = 0; c < H; ++c) ml[(size_t)i * H + c] += hl[(size_t)j * H + c];
            }
            if (deg[i])
                for (int c = 0; c < H; ++c) ml[(size_t)i * H + c] /= deg[i];
        }
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < H; ++c) {
                double s = b[c];
                for (int r = 0; r < H; ++r) s += hl[(size_t)i * H + r] * W[r * H + c];
                for (int r = 0; r < H; ++r) s += ml[(size_t)i * H + r] * W[(H + r) * H + c];
                zl[(size_t)i * H + c] = s;
                hn[(size_t)i * H + c] = hl[(size_t)i * H + c] + (s > 0 ? s : 0);
            }
    }
    const double* hL = h + (size_t)L * n * H;
    double loss = 0.0;
    double* dout = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        double o = bdec;
        for (int c = 0; c < H; ++c) o += wdec[c] * hL[(size_t)i * H + c];
        if (out) out[i] = o;
        double d = o - t[i];
        loss += d * d;
        dout[i] = dout_ext ? dout_ext[i] : 2.0 * d / n;
    }
    loss /= n;
    if (grad) {
        int np = ref_gnn_n_params(fin, H, L);
        memset(grad, 0, sizeof(double) * np);
        double* gWenc = grad;
        double* gbenc = gWenc + fin * H;
        double* glay = gbenc + H;
        double* gwdec = glay + L * (2 * H * H + H);
        double* g = (double*)malloc(sizeof(double) * (size_t)n * H);   /* dL/dh_l */
        double* dm = (double*)malloc(sizeof(double) * (size_t)n * H);
        for (int i = 0; i < n; ++i) {
            gwdec[H] += dout[i];
            for (int c = 0; c < H; ++c) {
                gwdec[c] += dout[i] * hL[(size_t)i * H + c];
                g[(size_t)i * H + c] = dout[i] * wdec[c];
            }
        }
        for (int l = L - 1; l >= 0; --l) {
            const double* W = lay + l * (2 * H * H + H);
            double* gW = glay + l * (2 * H * H + H);
            double* gb = gW + 2 * H * H;
            const double* hl = h + (size_t)l * n * H;
            const double* ml = m + (size_t)l * n * H;
            const double* zl = z + (size_t)(l + 1) * n * H;
            for (int i = 0; i < n; ++i) {
                double dz[512];
                for (int c = 0; c < H; ++c) dz[c] = zl[(size_t)i * H + c] > 0 ? g[(size_t)i * H + c] : 0.0;
                for (int c = 0; c < H; ++c) {
                    gb[c] += dz[c];
                    for (int r = 0; r < H; ++r) {
                        gW[r * H + c] += hl[(size_t)i * H + r] * dz[c];
                        gW[(H + r) * H + c] += ml[(size_t)i * H + r] * dz[c];
                    }
                }
                for (int r = 0; r < H; ++r) {
                    double sh = 0.0, sm = 0.0;
                    for (int c = 0; c < H; ++c) { sh += W[r * H + c] * dz[c]; sm += W[(H + r) * H + c] * dz[c]; }
                    g[(size_t)i * H + r] += sh;             /* residual path + self part */
                    dm[(size_t)i * H + r] = deg[i] ? sm / deg[i] : 0.0;
                }
            }
            /* m_i = mean_j h_j  ->  dh_j += sum_{i: j in N(i)} dm_i */
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < K; ++k) {
                    int j = nb[i * K + k];
                    if (j < 0) continue;
                    for (int c = 0; c < H; ++c) g[(size_t)j * H + c] += dm[(size_t)i * H + c];
                }
        }
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < H; ++c) {
                double d = z[(size_t)i * H + c] > 0 ? g[(size_t)i * H + c] : 0.0;
                gbenc[c] += d;
                for (int f = 0; f < fin; ++f) gWenc[f * H + c] += x[i * fin + f] * d;
            }
        free(g); free(dm);
    }
    free(h); free(m); free(z); free(deg); free(dout);
    return loss;
}
