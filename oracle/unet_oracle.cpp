// unet_oracle.cpp — TEST INFRASTRUCTURE ONLY: an FP64 CPU restatement of the warm-start
// network's forward pass, its backward pass and its optimiser, the checker of the HIP training
// kernels (diff-solver_amd/csrc/unet.hip, csrc/unet_train.hip). The reference has no learned
// part; the network is this build's addition, so the restatement follows the layer definitions
// below (which are also those of PyTorch's Conv3d / ConvTranspose3d / max_pool3d, against which
// tests/test_oracle_unet.py checks it) rather than reference source lines.
//
// Tensors are channels last: t[b][v][c] with v = x + nx (y + ny z).
//   conv3:  out[v][co] = bias[co] + sum_{ci, tap} w[co][ci][tap] in[v + off(tap)][ci], zero padding,
//           tap = (dx + 1) + 3 (dy + 1) + 9 (dz + 1)
//   tconv2: out[2 v + (kx, ky, kz)][co] = bias[co] + sum_ci w[ci][co][kx + 2 ky + 4 kz] in[v][ci]
//   pool:   out[v][c] = max over the 8 fine voxels 2 v + (dx, dy, dz); the backward routes to the
//           first maximum in the order t = dx + 2 dy + 4 dz
//   head:   out[v] = bias + sum_c w[c] in[v][c]
// The network (parameters in PyTorch state_dict order, per layer weight then bias):
//   s1 = R(c1b(R(c1a(x))));  s2 = R(c2b(R(c2a(P(s1)))));  c = R(cbb(R(cba(P(s2)))))
//   d2 = R(d2b(R(d2a([u2(c), s2]))));  d1 = R(d1b(R(d1a([u1(d2), s1]))));  out = head(d1)
// (R = ReLU, P = pool, [a, b] = channel concatenation). With bf16 = 1 every stored activation,
// activation gradient and the convolution weights are rounded to BF16 (nearest even) where the
// GPU stores BF16; sums stay FP64.
// The training loss of a residual-correction pair (r, e = A^-1 r) and a network output g (0 off
// the unknowns): L = 1 - <g, r>^2 / (<g, A g> <e, r>), the share of the A-norm error ||e||_A^2
// that the best multiple of g leaves.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using i64 = std::int64_t;

struct Dim {
    int nx, ny, nz;
    i64 n() const { return (i64)nx * ny * nz; }
    i64 at(int x, int y, int z) const { return x + (i64)nx * (y + (i64)ny * z); }
};

double bf16(double v) {   // FP64 -> FP32 -> BF16 (round to nearest even) -> FP64
    float f = (float)v;
    std::uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return v;   // inf / nan pass through
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

void round_all(std::vector<double>& v, bool on) {
    if (on)
        for (double& x : v) x = bf16(x);
}

void conv3_forward(int B, Dim d, int cin, int cout, const double* in, const double* w, const double* bias, bool relu, double* out) {
    const i64 n = d.n();
#pragma omp parallel for collapse(2)
    for (int b = 0; b < B; ++b)
        for (int z = 0; z < d.nz; ++z)
            for (int y = 0; y < d.ny; ++y)
                for (int x = 0; x < d.nx; ++x) {
                    double* o = out + ((i64)b * n + d.at(x, y, z)) * cout;
                    for (int co = 0; co < cout; ++co) o[co] = bias ? bias[co] : 0.0;
                    for (int tap = 0; tap < 27; ++tap) {
                        const int xx = x + tap % 3 - 1, yy = y + (tap / 3) % 3 - 1, zz = z + tap / 9 - 1;
                        if (xx < 0 || xx >= d.nx || yy < 0 || yy >= d.ny || zz < 0 || zz >= d.nz) continue;
                        const double* src = in + ((i64)b * n + d.at(xx, yy, zz)) * cin;
                        for (int co = 0; co < cout; ++co) {
                            double s = 0.0;
                            for (int ci = 0; ci < cin; ++ci) s += w[((i64)co * cin + ci) * 27 + tap] * src[ci];
                            o[co] += s;
                        }
                    }
                    if (relu)
                        for (int co = 0; co < cout; ++co) o[co] = std::max(o[co], 0.0);
                }
}

// din[v][ci] = sum_{tap, co} w[co][ci][tap] dout[v - off(tap)][co]
void conv3_dgrad(int B, Dim d, int cin, int cout, const double* dout, const double* w, double* din) {
    const i64 n = d.n();
#pragma omp parallel for collapse(2)
    for (int b = 0; b < B; ++b)
        for (int z = 0; z < d.nz; ++z)
            for (int y = 0; y < d.ny; ++y)
                for (int x = 0; x < d.nx; ++x) {
                    double* o = din + ((i64)b * n + d.at(x, y, z)) * cin;
                    for (int ci = 0; ci < cin; ++ci) o[ci] = 0.0;
                    for (int tap = 0; tap < 27; ++tap) {
                        const int xx = x - (tap % 3 - 1), yy = y - ((tap / 3) % 3 - 1), zz = z - (tap / 9 - 1);
                        if (xx < 0 || xx >= d.nx || yy < 0 || yy >= d.ny || zz < 0 || zz >= d.nz) continue;
                        const double* g = dout + ((i64)b * n + d.at(xx, yy, zz)) * cout;
                        for (int ci = 0; ci < cin; ++ci) {
                            double s = 0.0;
                            for (int co = 0; co < cout; ++co) s += w[((i64)co * cin + ci) * 27 + tap] * g[co];
                            o[ci] += s;
                        }
                    }
                }
}

// dw[co][ci][tap] = sum_v dout[v][co] in[v + off(tap)][ci]; db[co] = sum_v dout[v][co]
void conv3_wgrad(int B, Dim d, int cin, int cout, const double* in, const double* dout, double* dw, double* db) {
    const i64 n = d.n();
#pragma omp parallel for collapse(2)
    for (int co = 0; co < cout; ++co)
        for (int tap = 0; tap < 27; ++tap) {
            const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
            std::vector<double> acc(cin, 0.0);
            for (int b = 0; b < B; ++b)
                for (int z = 0; z < d.nz; ++z)
                    for (int y = 0; y < d.ny; ++y)
                        for (int x = 0; x < d.nx; ++x) {
                            const int xx = x + dx, yy = y + dy, zz = z + dz;
                            if (xx < 0 || xx >= d.nx || yy < 0 || yy >= d.ny || zz < 0 || zz >= d.nz) continue;
                            const double g = dout[((i64)b * n + d.at(x, y, z)) * cout + co];
                            const double* src = in + ((i64)b * n + d.at(xx, yy, zz)) * cin;
                            for (int ci = 0; ci < cin; ++ci) acc[ci] += g * src[ci];
                        }
            for (int ci = 0; ci < cin; ++ci) dw[((i64)co * cin + ci) * 27 + tap] = acc[ci];
        }
    if (db)
        for (int co = 0; co < cout; ++co) {
            double s = 0.0;
            for (i64 i = 0; i < (i64)B * n; ++i) s += dout[i * cout + co];
            db[co] = s;
        }
}

// coarse dc -> fine (2 dc)
void tconv2_forward(int B, Dim dc, int cin, int cout, const double* in, const double* w, const double* bias, double* out) {
    const Dim df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    const i64 nc = dc.n(), nf = df.n();
#pragma omp parallel for collapse(2)
    for (int b = 0; b < B; ++b)
        for (int z = 0; z < df.nz; ++z)
            for (int y = 0; y < df.ny; ++y)
                for (int x = 0; x < df.nx; ++x) {
                    const int t = (x & 1) + 2 * (y & 1) + 4 * (z & 1);
                    const double* src = in + ((i64)b * nc + dc.at(x >> 1, y >> 1, z >> 1)) * cin;
                    double* o = out + ((i64)b * nf + df.at(x, y, z)) * cout;
                    for (int co = 0; co < cout; ++co) {
                        double s = bias ? bias[co] : 0.0;
                        for (int ci = 0; ci < cin; ++ci) s += w[((i64)ci * cout + co) * 8 + t] * src[ci];
                        o[co] = s;
                    }
                }
}

void tconv2_dgrad(int B, Dim dc, int cin, int cout, const double* dout, const double* w, double* din) {
    const Dim df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    const i64 nc = dc.n(), nf = df.n();
#pragma omp parallel for collapse(2)
    for (int b = 0; b < B; ++b)
        for (int z = 0; z < dc.nz; ++z)
            for (int y = 0; y < dc.ny; ++y)
                for (int x = 0; x < dc.nx; ++x) {
                    double* o = din + ((i64)b * nc + dc.at(x, y, z)) * cin;
                    for (int ci = 0; ci < cin; ++ci) o[ci] = 0.0;
                    for (int t = 0; t < 8; ++t) {
                        const double* g = dout + ((i64)b * nf + df.at(2 * x + (t & 1), 2 * y + ((t >> 1) & 1), 2 * z + (t >> 2))) * cout;
                        for (int ci = 0; ci < cin; ++ci) {
                            double s = 0.0;
                            for (int co = 0; co < cout; ++co) s += w[((i64)ci * cout + co) * 8 + t] * g[co];
                            o[ci] += s;
                        }
                    }
                }
}

void tconv2_wgrad(int B, Dim dc, int cin, int cout, const double* in, const double* dout, double* dw, double* db) {
    const Dim df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    const i64 nc = dc.n(), nf = df.n();
#pragma omp parallel for collapse(2)
    for (int ci = 0; ci < cin; ++ci)
        for (int t = 0; t < 8; ++t) {
            std::vector<double> acc(cout, 0.0);
            for (int b = 0; b < B; ++b)
                for (int z = 0; z < dc.nz; ++z)
                    for (int y = 0; y < dc.ny; ++y)
                        for (int x = 0; x < dc.nx; ++x) {
                            const double a = in[((i64)b * nc + dc.at(x, y, z)) * cin + ci];
                            const double* g = dout + ((i64)b * nf + df.at(2 * x + (t & 1), 2 * y + ((t >> 1) & 1), 2 * z + (t >> 2))) * cout;
                            for (int co = 0; co < cout; ++co) acc[co] += a * g[co];
                        }
            for (int co = 0; co < cout; ++co) dw[((i64)ci * cout + co) * 8 + t] = acc[co];
        }
    if (db)
        for (int co = 0; co < cout; ++co) {
            double s = 0.0;
            for (i64 i = 0; i < (i64)B * nf; ++i) s += dout[i * cout + co];
            db[co] = s;
        }
}

void maxpool_forward(int B, Dim dc, int c, const double* in, double* out) {
    const Dim df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    const i64 nc = dc.n(), nf = df.n();
    for (int b = 0; b < B; ++b)
        for (int z = 0; z < dc.nz; ++z)
            for (int y = 0; y < dc.ny; ++y)
                for (int x = 0; x < dc.nx; ++x)
                    for (int ch = 0; ch < c; ++ch) {
                        double m = -INFINITY;
                        for (int t = 0; t < 8; ++t)
                            m = std::max(m, in[((i64)b * nf + df.at(2 * x + (t & 1), 2 * y + ((t >> 1) & 1), 2 * z + (t >> 2))) * c + ch]);
                        out[((i64)b * nc + dc.at(x, y, z)) * c + ch] = m;
                    }
}

// din += the coarse gradient routed to the first maximum of each group
void maxpool_backward(int B, Dim dc, int c, const double* in, const double* dout, double* din) {
    const Dim df{2 * dc.nx, 2 * dc.ny, 2 * dc.nz};
    const i64 nc = dc.n(), nf = df.n();
    for (int b = 0; b < B; ++b)
        for (int z = 0; z < dc.nz; ++z)
            for (int y = 0; y < dc.ny; ++y)
                for (int x = 0; x < dc.nx; ++x)
                    for (int ch = 0; ch < c; ++ch) {
                        double m = -INFINITY;
                        i64 arg = -1;
                        for (int t = 0; t < 8; ++t) {
                            const i64 v = df.at(2 * x + (t & 1), 2 * y + ((t >> 1) & 1), 2 * z + (t >> 2));
                            const double val = in[((i64)b * nf + v) * c + ch];
                            if (val > m) { m = val; arg = v; }
                        }
                        din[((i64)b * nf + arg) * c + ch] += dout[((i64)b * nc + dc.at(x, y, z)) * c + ch];
                    }
}

// ---- the network ------------------------------------------------------------------------------
struct Layer {
    int kind;   // 0 conv3, 1 tconv2, 2 head
    int cin, cout;
};
constexpr Layer kLayers[13] = {
    {0, 2, 16},  {0, 16, 16}, {0, 16, 32}, {0, 32, 32}, {0, 32, 64}, {0, 64, 64}, {1, 64, 32},
    {0, 64, 32}, {0, 32, 32}, {1, 32, 16}, {0, 32, 16}, {0, 16, 16}, {2, 16, 1},
};

i64 layer_weights(const Layer& L) { return (i64)L.cin * L.cout * (L.kind == 0 ? 27 : L.kind == 1 ? 8 : 1); }

struct Params {
    const double* w[13];
    const double* b[13];
    std::vector<double> rounded[13];   // BF16-rounded convolution weights (bf16 mode)
    Params(const double* flat, bool bf) {
        for (int k = 0; k < 13; ++k) {
            const i64 nw = layer_weights(kLayers[k]);
            w[k] = flat;
            b[k] = flat + nw;
            if (bf && kLayers[k].kind != 2) {
                rounded[k].assign(flat, flat + nw);
                round_all(rounded[k], true);
                w[k] = rounded[k].data();
            }
            flat += nw + kLayers[k].cout;
        }
    }
};

struct Acts {
    std::vector<double> x, a1, s1, p1, a2, s2, p2, a3, c, u2, cat2, a4, d2, u1, cat1, a5, d1;
};

void concat(int B, i64 n, const std::vector<double>& a, int ca, const std::vector<double>& b, int cb, std::vector<double>& out) {
    out.assign((size_t)B * n * (ca + cb), 0.0);
    for (i64 i = 0; i < (i64)B * n; ++i) {
        for (int c = 0; c < ca; ++c) out[i * (ca + cb) + c] = a[i * ca + c];
        for (int c = 0; c < cb; ++c) out[i * (ca + cb) + ca + c] = b[i * cb + c];
    }
}

void network_forward(int B, Dim d, const Params& P, const double* in2, bool bf, Acts& A, double* out) {
    const Dim d2{d.nx / 2, d.ny / 2, d.nz / 2}, d3{d.nx / 4, d.ny / 4, d.nz / 4};
    const i64 n1 = d.n(), n2 = d2.n(), n3 = d3.n();
    auto conv = [&](int k, Dim dd, const std::vector<double>& in, std::vector<double>& o) {
        o.assign((size_t)B * dd.n() * kLayers[k].cout, 0.0);
        conv3_forward(B, dd, kLayers[k].cin, kLayers[k].cout, in.data(), P.w[k], P.b[k], true, o.data());
        round_all(o, bf);
    };
    A.x.assign(in2, in2 + (size_t)B * n1 * 2);
    round_all(A.x, bf);
    conv(0, d, A.x, A.a1);
    conv(1, d, A.a1, A.s1);
    A.p1.assign((size_t)B * n2 * 16, 0.0);
    maxpool_forward(B, d2, 16, A.s1.data(), A.p1.data());
    conv(2, d2, A.p1, A.a2);
    conv(3, d2, A.a2, A.s2);
    A.p2.assign((size_t)B * n3 * 32, 0.0);
    maxpool_forward(B, d3, 32, A.s2.data(), A.p2.data());
    conv(4, d3, A.p2, A.a3);
    conv(5, d3, A.a3, A.c);
    A.u2.assign((size_t)B * n2 * 32, 0.0);
    tconv2_forward(B, d3, 64, 32, A.c.data(), P.w[6], P.b[6], A.u2.data());
    round_all(A.u2, bf);
    concat(B, n2, A.u2, 32, A.s2, 32, A.cat2);
    conv(7, d2, A.cat2, A.a4);
    conv(8, d2, A.a4, A.d2);
    A.u1.assign((size_t)B * n1 * 16, 0.0);
    tconv2_forward(B, d2, 32, 16, A.d2.data(), P.w[9], P.b[9], A.u1.data());
    round_all(A.u1, bf);
    concat(B, n1, A.u1, 16, A.s1, 16, A.cat1);
    conv(10, d, A.cat1, A.a5);
    conv(11, d, A.a5, A.d1);
    for (i64 i = 0; i < (i64)B * n1; ++i) {
        double s = P.b[12][0];
        for (int c = 0; c < 16; ++c) s += P.w[12][c] * A.d1[i * 16 + c];
        out[i] = s;
    }
}

// gradient of a ReLU output y = R(z) with respect to z, from the gradient with respect to y
void relu_mask(std::vector<double>& g, const std::vector<double>& y, bool bf) {
    for (size_t i = 0; i < g.size(); ++i) g[i] = y[i] > 0.0 ? (bf ? bf16(g[i]) : g[i]) : 0.0;
}

void network_backward(int B, Dim d, const Params& P, const Acts& A, const double* dout, bool bf, double* grad) {
    const Dim d2{d.nx / 2, d.ny / 2, d.nz / 2}, d3{d.nx / 4, d.ny / 4, d.nz / 4};
    const i64 n1 = d.n(), n2 = d2.n(), n3 = d3.n();
    double* gw[13];
    double* gb[13];
    {
        double* p = grad;
        for (int k = 0; k < 13; ++k) {
            gw[k] = p;
            gb[k] = p + layer_weights(kLayers[k]);
            p += layer_weights(kLayers[k]) + kLayers[k].cout;
        }
    }
    // head
    std::vector<double> g_d1((size_t)B * n1 * 16);
    for (int c = 0; c < 16; ++c) gw[12][c] = 0.0;
    gb[12][0] = 0.0;
    for (i64 i = 0; i < (i64)B * n1; ++i) {
        gb[12][0] += dout[i];
        for (int c = 0; c < 16; ++c) {
            gw[12][c] += dout[i] * A.d1[i * 16 + c];
            g_d1[i * 16 + c] = dout[i] * P.w[12][c];
        }
    }
    relu_mask(g_d1, A.d1, bf);   // now the gradient of d1b's pre-activation
    // conv layer k: given dz (pre-activation gradient, already masked), weight gradients, and the
    // gradient of its input (unmasked)
    auto conv_back = [&](int k, Dim dd, const std::vector<double>& in, const std::vector<double>& dz, std::vector<double>* din) {
        conv3_wgrad(B, dd, kLayers[k].cin, kLayers[k].cout, in.data(), dz.data(), gw[k], gb[k]);
        if (din) {
            din->assign((size_t)B * dd.n() * kLayers[k].cin, 0.0);
            conv3_dgrad(B, dd, kLayers[k].cin, kLayers[k].cout, dz.data(), P.w[k], din->data());
        }
    };
    std::vector<double> g_a5, g_cat1, g_u1((size_t)B * n1 * 16), g_s1((size_t)B * n1 * 16);
    conv_back(11, d, A.a5, g_d1, &g_a5);
    relu_mask(g_a5, A.a5, bf);
    conv_back(10, d, A.cat1, g_a5, &g_cat1);
    for (i64 i = 0; i < (i64)B * n1; ++i)
        for (int c = 0; c < 16; ++c) {
            g_u1[i * 16 + c] = bf ? bf16(g_cat1[i * 32 + c]) : g_cat1[i * 32 + c];
            g_s1[i * 16 + c] = g_cat1[i * 32 + 16 + c];
        }
    // up1: tconv from d2
    std::vector<double> g_d2((size_t)B * n2 * 32);
    tconv2_wgrad(B, d2, 32, 16, A.d2.data(), g_u1.data(), gw[9], gb[9]);
    tconv2_dgrad(B, d2, 32, 16, g_u1.data(), P.w[9], g_d2.data());
    relu_mask(g_d2, A.d2, bf);
    std::vector<double> g_a4, g_cat2, g_u2((size_t)B * n2 * 32), g_s2((size_t)B * n2 * 32);
    conv_back(8, d2, A.a4, g_d2, &g_a4);
    relu_mask(g_a4, A.a4, bf);
    conv_back(7, d2, A.cat2, g_a4, &g_cat2);
    for (i64 i = 0; i < (i64)B * n2; ++i)
        for (int c = 0; c < 32; ++c) {
            g_u2[i * 32 + c] = bf ? bf16(g_cat2[i * 64 + c]) : g_cat2[i * 64 + c];
            g_s2[i * 32 + c] = g_cat2[i * 64 + 32 + c];
        }
    std::vector<double> g_c((size_t)B * n3 * 64);
    tconv2_wgrad(B, d3, 64, 32, A.c.data(), g_u2.data(), gw[6], gb[6]);
    tconv2_dgrad(B, d3, 64, 32, g_u2.data(), P.w[6], g_c.data());
    relu_mask(g_c, A.c, bf);
    std::vector<double> g_a3, g_p2;
    conv_back(5, d3, A.a3, g_c, &g_a3);
    relu_mask(g_a3, A.a3, bf);
    conv_back(4, d3, A.p2, g_a3, &g_p2);
    if (bf) round_all(g_p2, true);
    maxpool_backward(B, d3, 32, A.s2.data(), g_p2.data(), g_s2.data());   // skip + pool paths
    relu_mask(g_s2, A.s2, bf);
    std::vector<double> g_a2, g_p1;
    conv_back(3, d2, A.a2, g_s2, &g_a2);
    relu_mask(g_a2, A.a2, bf);
    conv_back(2, d2, A.p1, g_a2, &g_p1);
    if (bf) round_all(g_p1, true);
    maxpool_backward(B, d2, 16, A.s1.data(), g_p1.data(), g_s1.data());
    relu_mask(g_s1, A.s1, bf);
    std::vector<double> g_a1;
    conv_back(1, d, A.a1, g_s1, &g_a1);
    relu_mask(g_a1, A.a1, bf);
    conv_back(0, d, A.x, g_a1, nullptr);
}

// A g on the unknowns (code = diagonal count, 0 off the unknowns; a neighbour couples when it is
// an unknown; periodic bit a wraps axis a)
void apply_matrix(Dim d, const std::uint8_t* code, unsigned per, const double* g, double* out) {
#pragma omp parallel for
    for (int z = 0; z < d.nz; ++z)
        for (int y = 0; y < d.ny; ++y)
            for (int x = 0; x < d.nx; ++x) {
                const i64 i = d.at(x, y, z);
                if (!code[i]) { out[i] = 0.0; continue; }
                double s = code[i] * g[i];
                const int c[3]{x, y, z}, n[3]{d.nx, d.ny, d.nz};
                for (int a = 0; a < 3; ++a)
                    for (int dir = -1; dir <= 1; dir += 2) {
                        int q[3]{c[0], c[1], c[2]};
                        q[a] += dir;
                        if (q[a] < 0 || q[a] >= n[a]) {
                            if (!(per >> a & 1u)) continue;
                            q[a] = q[a] < 0 ? n[a] - 1 : 0;
                        }
                        const i64 j = d.at(q[0], q[1], q[2]);
                        if (code[j]) s -= g[j];
                    }
                out[i] = s;
            }
}

}  // namespace

extern "C" {

void unet_conv3_forward(int B, int nx, int ny, int nz, int cin, int cout, const double* in, const double* w, const double* bias, int relu,
                        double* out) {
    conv3_forward(B, Dim{nx, ny, nz}, cin, cout, in, w, bias, relu != 0, out);
}
void unet_conv3_dgrad(int B, int nx, int ny, int nz, int cin, int cout, const double* dout, const double* w, double* din) {
    conv3_dgrad(B, Dim{nx, ny, nz}, cin, cout, dout, w, din);
}
void unet_conv3_wgrad(int B, int nx, int ny, int nz, int cin, int cout, const double* in, const double* dout, double* dw, double* db) {
    conv3_wgrad(B, Dim{nx, ny, nz}, cin, cout, in, dout, dw, db);
}
void unet_tconv2_forward(int B, int nx, int ny, int nz, int cin, int cout, const double* in, const double* w, const double* bias,
                         double* out) {
    tconv2_forward(B, Dim{nx, ny, nz}, cin, cout, in, w, bias, out);
}
void unet_tconv2_dgrad(int B, int nx, int ny, int nz, int cin, int cout, const double* dout, const double* w, double* din) {
    tconv2_dgrad(B, Dim{nx, ny, nz}, cin, cout, dout, w, din);
}
void unet_tconv2_wgrad(int B, int nx, int ny, int nz, int cin, int cout, const double* in, const double* dout, double* dw, double* db) {
    tconv2_wgrad(B, Dim{nx, ny, nz}, cin, cout, in, dout, dw, db);
}
void unet_maxpool_forward(int B, int nx, int ny, int nz, int c, const double* in, double* out) {
    maxpool_forward(B, Dim{nx, ny, nz}, c, in, out);
}
void unet_maxpool_backward(int B, int nx, int ny, int nz, int c, const double* in, const double* dout, double* din) {
    maxpool_backward(B, Dim{nx, ny, nz}, c, in, dout, din);
}
std::int64_t unet_parameter_count(void) {
    std::int64_t n = 0;
    for (const Layer& L : kLayers) n += layer_weights(L) + L.cout;
    return n;
}

// The whole network on B samples of the fine grid (nx, ny, nz): in2 [B][N][2] -> out [B][N]; with
// dout != null also the parameter gradient (flat, state_dict order) of <dout, out>.
void unet_network(int B, int nx, int ny, int nz, const double* params, const double* in2, int bf16_mode, double* out, const double* dout,
                  double* grad) {
    const Dim d{nx, ny, nz};
    const Params P(params, bf16_mode != 0);
    Acts A;
    network_forward(B, d, P, in2, bf16_mode != 0, A, out);
    if (dout && grad) network_backward(B, d, P, A, dout, bf16_mode != 0, grad);
}

// Energy loss of B pairs and its gradient with respect to g (ddelta, same shape), per sample:
// loss[b] = 1 - q^2 / (D E), q = <g, r>, D = <g, A g>, E = <e, r>.
void unet_energy_loss(int B, int nx, int ny, int nz, const std::uint8_t* code, unsigned periodic, const double* g, const double* r,
                      const double* e, double* loss, double* dg) {
    const Dim d{nx, ny, nz};
    const i64 n = d.n();
    std::vector<double> ag(n);
    for (int b = 0; b < B; ++b) {
        const double* gb = g + (i64)b * n;
        const double* rb = r + (i64)b * n;
        const double* eb = e + (i64)b * n;
        apply_matrix(d, code, periodic, gb, ag.data());
        double q = 0.0, D = 0.0, E = 0.0;
        for (i64 i = 0; i < n; ++i) {
            if (!code[i]) continue;
            q += gb[i] * rb[i];
            D += gb[i] * ag[i];
            E += eb[i] * rb[i];
        }
        loss[b] = 1.0 - q * q / (D * E);
        if (dg) {
            // dL/dg = -(2 q / (D E)) (r - (q / D) A g)
            const double s = -2.0 * q / (D * E);
            for (i64 i = 0; i < n; ++i) dg[(i64)b * n + i] = code[i] ? s * (rb[i] - q / D * ag[i]) : 0.0;
        }
    }
}

// L2 loss of B samples and its gradient with respect to g, per sample: with r = A e,
// gamma = <g, r> / <g, A g> (the step the solver applies) and f = gamma g - e:
// loss[b] = ||f||^2 / ||e||^2, dL/dg = (2 <f, g> (r - 2 gamma A g) / <g, A g> + 2 gamma f) / ||e||^2.
void unet_l2_loss(int B, int nx, int ny, int nz, const std::uint8_t* code, unsigned periodic, const double* g, const double* e,
                  double* loss, double* dg) {
    const Dim d{nx, ny, nz};
    const i64 n = d.n();
    std::vector<double> ag(n), r(n);
    for (int b = 0; b < B; ++b) {
        const double* gb = g + (i64)b * n;
        const double* eb = e + (i64)b * n;
        apply_matrix(d, code, periodic, gb, ag.data());
        apply_matrix(d, code, periodic, eb, r.data());
        double q = 0.0, D = 0.0, GG = 0.0, P = 0.0, E2 = 0.0;
        for (i64 i = 0; i < n; ++i) {
            if (!code[i]) continue;
            q += gb[i] * r[i];
            D += gb[i] * ag[i];
            GG += gb[i] * gb[i];
            P += gb[i] * eb[i];
            E2 += eb[i] * eb[i];
        }
        const double gamma = q / D;
        loss[b] = (gamma * gamma * GG - 2.0 * gamma * P + E2) / E2;
        if (dg) {
            const double fg = gamma * GG - P;
            for (i64 i = 0; i < n; ++i) {
                const double f = gamma * gb[i] - eb[i];
                dg[(i64)b * n + i] = code[i] ? (2.0 * fg * (r[i] - 2.0 * gamma * ag[i]) / D + 2.0 * gamma * f) / E2 : 0.0;
            }
        }
    }
}

// Adam (Kingma & Ba, bias-corrected): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps); step t counts from 1.
void unet_adam(std::int64_t n, double* p, const double* g, double* m, double* v, int t, double lr, double b1, double b2, double eps) {
    const double c1 = 1.0 - std::pow(b1, t), c2 = 1.0 - std::pow(b2, t);
    for (std::int64_t i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        p[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
    }
}

}  // extern "C"
