// CPU oracle of the learned warm-start network — TEST INFRASTRUCTURE ONLY.
//
// The reference has no learned component; this is a plain restatement of the
// network the product builds (3-level 3-D U-Net, see include/xr.h
// xr_unet_create) in FP64 direct convolutions, with its analytic backward
// pass, so that the product's BF16 MFMA kernels can be checked against it.
// With emulate_bf16 = 1 the values are rounded to BF16 (round to nearest
// even) at exactly the points where the product stores BF16: packed weights,
// the scaled input, every activation, pooled values and every activation
// gradient; sums stay FP64 (the product sums in FP32).  With emulate_bf16 = 0
// everything is FP64: that mode is pinned by finite differences of the loss
// (tests/test_oracle_unet.py).
//
// Layout: volumes v = p2 + D2*(p1 + D1*p0); activations act[v*C + c].
// Parameters: flat, per layer W[Cout][27][Cin] then b[Cout]; head w[C], b.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline double bf16r(double x) {
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

struct Net {
    int C;
    int io[10][2];
    int64_t w_off[10], b_off[10], head, n;
    explicit Net(int c) : C(c) {
        const int t[10][2] = {{8, c},         {c, c},         {c, 2 * c},     {2 * c, 2 * c}, {2 * c, 4 * c},
                              {4 * c, 4 * c}, {6 * c, 2 * c}, {2 * c, 2 * c}, {3 * c, c},     {c, c}};
        int64_t off = 0;
        for (int i = 0; i < 10; i++) {
            io[i][0] = t[i][0];
            io[i][1] = t[i][1];
            w_off[i] = off;
            off += (int64_t)t[i][1] * 27 * t[i][0];
            b_off[i] = off;
            off += t[i][1];
        }
        head = off;
        n = off + c + 1;
    }
};

struct Dims {
    int d0, d1, d2;
    int64_t V() const { return (int64_t)d0 * d1 * d2; }
    Dims half() const { return {d0 / 2, d1 / 2, d2 / 2}; }
};

typedef std::vector<double> Buf;

// Tensor view: data + row length (ld) + channel offset.
struct T {
    double* p;
    int ld, off;
    double& at(int64_t v, int c) const { return p[v * ld + off + c]; }
};

void conv_fwd(const T& in, int Cin, const Dims& D, const double* W, const double* b, const T& out, int Cout,
              bool relu, bool em) {
    for (int p0 = 0; p0 < D.d0; p0++)
        for (int p1 = 0; p1 < D.d1; p1++)
            for (int p2 = 0; p2 < D.d2; p2++) {
                const int64_t v = ((int64_t)p0 * D.d1 + p1) * D.d2 + p2;
                for (int co = 0; co < Cout; co++) {
                    double s = b ? b[co] : 0.0;
                    for (int tap = 0; tap < 27; tap++) {
                        const int n0 = p0 + tap / 9 - 1, n1 = p1 + (tap / 3) % 3 - 1, n2 = p2 + tap % 3 - 1;
                        if (n0 < 0 || n0 >= D.d0 || n1 < 0 || n1 >= D.d1 || n2 < 0 || n2 >= D.d2) continue;
                        const int64_t nv = ((int64_t)n0 * D.d1 + n1) * D.d2 + n2;
                        const double* w = W + ((int64_t)co * 27 + tap) * Cin;
                        for (int ci = 0; ci < Cin; ci++) s += w[ci] * in.at(nv, ci);
                    }
                    if (relu && s < 0) s = 0;
                    out.at(v, co) = em ? bf16r(s) : s;
                }
            }
}

// gin[v][ci] = sum_{tap,co} W[co][tap][ci] gout[v - off(tap)][co], masked, (+ old).
void conv_dgrad(const T& gout, int Cout, const Dims& D, const double* W, const T& gin, int Cin, const T* mask,
                bool em) {
    for (int p0 = 0; p0 < D.d0; p0++)
        for (int p1 = 0; p1 < D.d1; p1++)
            for (int p2 = 0; p2 < D.d2; p2++) {
                const int64_t v = ((int64_t)p0 * D.d1 + p1) * D.d2 + p2;
                for (int ci = 0; ci < Cin; ci++) {
                    double s = 0.0;
                    for (int tap = 0; tap < 27; tap++) {
                        // output voxel o with o + off(tap) = v
                        const int o0 = p0 - (tap / 9 - 1), o1 = p1 - ((tap / 3) % 3 - 1), o2 = p2 - (tap % 3 - 1);
                        if (o0 < 0 || o0 >= D.d0 || o1 < 0 || o1 >= D.d1 || o2 < 0 || o2 >= D.d2) continue;
                        const int64_t ov = ((int64_t)o0 * D.d1 + o1) * D.d2 + o2;
                        for (int co = 0; co < Cout; co++) s += W[((int64_t)co * 27 + tap) * Cin + ci] * gout.at(ov, co);
                    }
                    if (mask && !(mask->at(v, ci) > 0)) s = 0;
                    gin.at(v, ci) = em ? bf16r(s) : s;
                }
            }
}

void conv_wgrad(const T& in, int Cin, const T& gout, int Cout, const Dims& D, double* dW, double* db) {
    for (int p0 = 0; p0 < D.d0; p0++)
        for (int p1 = 0; p1 < D.d1; p1++)
            for (int p2 = 0; p2 < D.d2; p2++) {
                const int64_t v = ((int64_t)p0 * D.d1 + p1) * D.d2 + p2;
                for (int co = 0; co < Cout; co++) {
                    const double g = gout.at(v, co);
                    if (g == 0.0) continue;
                    db[co] += g;
                    for (int tap = 0; tap < 27; tap++) {
                        const int n0 = p0 + tap / 9 - 1, n1 = p1 + (tap / 3) % 3 - 1, n2 = p2 + tap % 3 - 1;
                        if (n0 < 0 || n0 >= D.d0 || n1 < 0 || n1 >= D.d1 || n2 < 0 || n2 >= D.d2) continue;
                        const int64_t nv = ((int64_t)n0 * D.d1 + n1) * D.d2 + n2;
                        double* w = dW + ((int64_t)co * 27 + tap) * Cin;
                        for (int ci = 0; ci < Cin; ci++) w[ci] += g * in.at(nv, ci);
                    }
                }
            }
}

inline int64_t child(const Dims& D, int64_t p, int c) {
    const Dims H = D.half();
    const int q2 = (int)(p % H.d2), q1 = (int)((p / H.d2) % H.d1), q0 = (int)(p / ((int64_t)H.d2 * H.d1));
    return ((int64_t)(2 * q0 + (c >> 2)) * D.d1 + (2 * q1 + ((c >> 1) & 1))) * D.d2 + (2 * q2 + (c & 1));
}
inline int64_t parent(const Dims& D, int64_t v) {
    const int p2 = (int)(v % D.d2), p1 = (int)((v / D.d2) % D.d1), p0 = (int)(v / ((int64_t)D.d2 * D.d1));
    return ((int64_t)(p0 / 2) * (D.d1 / 2) + p1 / 2) * (D.d2 / 2) + p2 / 2;
}

void pool(const T& in, int C, const Dims& D, const T& out, bool em) {
    const int64_t Vh = D.half().V();
    for (int64_t p = 0; p < Vh; p++)
        for (int c = 0; c < C; c++) {
            double s = 0;
            for (int k = 0; k < 8; k++) s += in.at(child(D, p, k), c);
            out.at(p, c) = em ? bf16r(0.125 * s) : 0.125 * s;
        }
}

void upsample(const T& low, int C, const Dims& D, const T& out) {
    for (int64_t v = 0; v < D.V(); v++)
        for (int c = 0; c < C; c++) out.at(v, c) = low.at(parent(D, v), c);
}

struct State {
    Net net;
    Dims D1, D2, D3;
    double s;
    bool em;
    std::vector<double> P;  // parameters (BF16-rounded weights when em)
    Buf x8, a1, cat1, p1, b1, cat2, p2, c1, c2, d1, d2, e1, e2;
    State(int C, Dims d, double scale, bool e) : net(C), D1(d), D2(d.half()), D3(d.half().half()), s(scale), em(e) {}
    T t(Buf& b, int ld, int off = 0) { return T{b.data(), ld, off}; }
};

void forward(State& S, const float* x, double* y) {
    const int C = S.net.C;
    const int64_t V1 = S.D1.V(), V2 = S.D2.V(), V3 = S.D3.V();
    S.x8.assign(V1 * 8, 0.0);
    S.a1.assign(V1 * C, 0);
    S.cat1.assign(V1 * 3 * C, 0);
    S.p1.assign(V2 * C, 0);
    S.b1.assign(V2 * 2 * C, 0);
    S.cat2.assign(V2 * 6 * C, 0);
    S.p2.assign(V3 * 2 * C, 0);
    S.c1.assign(V3 * 4 * C, 0);
    S.c2.assign(V3 * 4 * C, 0);
    S.d1.assign(V2 * 2 * C, 0);
    S.d2.assign(V2 * 2 * C, 0);
    S.e1.assign(V1 * C, 0);
    S.e2.assign(V1 * C, 0);
    const float inv = 1.0f / (float)S.s;
    for (int64_t v = 0; v < V1; v++) S.x8[v * 8] = S.em ? bf16r((double)(x[v] * inv)) : (double)x[v] / S.s;
    auto L = [&](int i, const T& in, const Dims& D, const T& out) {
        conv_fwd(in, S.net.io[i][0], D, S.P.data() + S.net.w_off[i], S.P.data() + S.net.b_off[i], out, S.net.io[i][1],
                 true, S.em);
    };
    L(0, S.t(S.x8, 8), S.D1, S.t(S.a1, C));
    L(1, S.t(S.a1, C), S.D1, S.t(S.cat1, 3 * C, 2 * C));
    pool(S.t(S.cat1, 3 * C, 2 * C), C, S.D1, S.t(S.p1, C), S.em);
    L(2, S.t(S.p1, C), S.D2, S.t(S.b1, 2 * C));
    L(3, S.t(S.b1, 2 * C), S.D2, S.t(S.cat2, 6 * C, 4 * C));
    pool(S.t(S.cat2, 6 * C, 4 * C), 2 * C, S.D2, S.t(S.p2, 2 * C), S.em);
    L(4, S.t(S.p2, 2 * C), S.D3, S.t(S.c1, 4 * C));
    L(5, S.t(S.c1, 4 * C), S.D3, S.t(S.c2, 4 * C));
    upsample(S.t(S.c2, 4 * C), 4 * C, S.D2, S.t(S.cat2, 6 * C, 0));
    L(6, S.t(S.cat2, 6 * C), S.D2, S.t(S.d1, 2 * C));
    L(7, S.t(S.d1, 2 * C), S.D2, S.t(S.d2, 2 * C));
    upsample(S.t(S.d2, 2 * C), 2 * C, S.D1, S.t(S.cat1, 3 * C, 0));
    L(8, S.t(S.cat1, 3 * C), S.D1, S.t(S.e1, C));
    L(9, S.t(S.e1, C), S.D1, S.t(S.e2, C));
    const double* w = S.P.data() + S.net.head;
    for (int64_t v = 0; v < V1; v++) {
        double f = w[C];
        for (int c = 0; c < C; c++) f += w[c] * S.e2[v * C + c];
        y[v] = f;  // network output in normalised units (before the residual)
    }
}

State make_state(const float* params, int C, int d0, int d1, int d2, float scale, int em) {
    State S(C, Dims{d0, d1, d2}, scale, em != 0);
    S.P.assign(params, params + S.net.n);
    if (S.em)
        for (int i = 0; i < 10; i++)
            for (int64_t j = 0; j < (int64_t)S.net.io[i][1] * 27 * S.net.io[i][0]; j++)
                S.P[S.net.w_off[i] + j] = bf16r(S.P[S.net.w_off[i] + j]);
    return S;
}

}  // namespace

extern "C" {

int64_t orc_unet_num_params(int C) { return Net(C).n; }

// One 3x3x3 convolution (FP64), in[v][Cin] -> out[v][Cout]: the building block
// of the network, exposed so that its tap order and padding can be pinned
// against a library correlation.
void orc_conv3(const float* in, int Cin, int d0, int d1, int d2, const float* W, const float* b, int Cout, int relu,
               double* out) {
    const Dims D{d0, d1, d2};
    std::vector<double> x(in, in + D.V() * Cin), w(W, W + (int64_t)Cout * 27 * Cin), bb(b, b + Cout);
    conv_fwd(T{x.data(), Cin, 0}, Cin, D, w.data(), bb.data(), T{out, Cout, 0}, Cout, relu != 0, false);
}

// y = x + s*net(x/s) (FP64 result rounded to float).
void orc_unet_forward(const float* params, int C, const float* x, int d0, int d1, int d2, float scale,
                      int emulate_bf16, float* y) {
    State S = make_state(params, C, d0, d1, d2, scale, emulate_bf16);
    std::vector<double> f(S.D1.V());
    forward(S, x, f.data());
    for (int64_t v = 0; v < S.D1.V(); v++) y[v] = (float)((double)x[v] + S.s * f[v]);
}

// loss = mean((x/s + net(x/s)) - t/s)^2 and d loss / d params.
double orc_unet_loss_grad(const float* params, int C, const float* x, const float* t, int d0, int d1, int d2,
                          float scale, int emulate_bf16, double* grads) {
    State S = make_state(params, C, d0, d1, d2, scale, emulate_bf16);
    const Net& N = S.net;
    const bool em = S.em;
    const int64_t V1 = S.D1.V(), V2 = S.D2.V(), V3 = S.D3.V();
    std::vector<double> f(V1);
    forward(S, x, f.data());
    std::fill(grads, grads + N.n, 0.0);
    const float inv = 1.0f / (float)S.s;
    std::vector<double> g(V1);
    double loss = 0;
    for (int64_t v = 0; v < V1; v++) {
        double r;
        if (em) {
            // the product forms r = (x*inv + f) - t*inv in FP32
            const float rf = ((float)(x[v] * inv) + (float)f[v]) - t[v] * inv;
            r = rf;
            g[v] = (double)(2.0f * rf / (float)V1);
        } else {
            r = (double)x[v] / S.s + f[v] - (double)t[v] / S.s;
            g[v] = 2.0 * r / (double)V1;
        }
        loss += r * r;
    }
    Buf g_e2(V1 * C), g_e1(V1 * C), g_cat1(V1 * 3 * C), g_d2(V2 * 2 * C), g_d1(V2 * 2 * C), g_cat2(V2 * 6 * C),
        g_c2(V3 * 4 * C), g_c1(V3 * 4 * C), g_p2(V3 * 2 * C), g_b1(V2 * 2 * C), g_p1(V2 * C), g_a1(V1 * C);
    const double* wh = S.P.data() + N.head;
    double* gh = grads + N.head;
    for (int64_t v = 0; v < V1; v++) {
        for (int c = 0; c < C; c++) {
            const double e = S.e2[v * C + c];
            const double val = e > 0 ? g[v] * wh[c] : 0.0;
            g_e2[v * C + c] = em ? bf16r(val) : val;
            gh[c] += g[v] * e;
        }
        gh[C] += g[v];
    }
    auto T_ = [](Buf& b, int ld, int off = 0) { return T{b.data(), ld, off}; };
    auto bwd = [&](int i, const T& in, const T& gout, const Dims& D, const T* gin, const T* mask) {
        conv_wgrad(in, N.io[i][0], gout, N.io[i][1], D, grads + N.w_off[i], grads + N.b_off[i]);
        if (gin) conv_dgrad(gout, N.io[i][1], D, S.P.data() + N.w_off[i], *gin, N.io[i][0], mask, em);
    };
    T te1 = T_(S.e1, C), tcat1 = T_(S.cat1, 3 * C), td1 = T_(S.d1, 2 * C), tcat2 = T_(S.cat2, 6 * C),
      tc1 = T_(S.c1, 4 * C), tb1 = T_(S.b1, 2 * C), ta1 = T_(S.a1, C);
    T ge1 = T_(g_e1, C), gcat1 = T_(g_cat1, 3 * C), gd1 = T_(g_d1, 2 * C), gcat2 = T_(g_cat2, 6 * C),
      gc1 = T_(g_c1, 4 * C), gp2 = T_(g_p2, 2 * C), gb1 = T_(g_b1, 2 * C), gp1 = T_(g_p1, C), ga1 = T_(g_a1, C);
    bwd(9, te1, T_(g_e2, C), S.D1, &ge1, &te1);
    bwd(8, tcat1, ge1, S.D1, &gcat1, &tcat1);
    // upsample backward: g_d2[p] = sum of children of g_cat1[:, 0:2C)
    for (int64_t p = 0; p < V2; p++)
        for (int c = 0; c < 2 * C; c++) {
            double s = 0;
            for (int k = 0; k < 8; k++) s += g_cat1[child(S.D1, p, k) * 3 * C + c];
            g_d2[p * 2 * C + c] = em ? bf16r(s) : s;
        }
    bwd(7, td1, T_(g_d2, 2 * C), S.D2, &gd1, &td1);
    bwd(6, tcat2, gd1, S.D2, &gcat2, &tcat2);
    for (int64_t p = 0; p < V3; p++)
        for (int c = 0; c < 4 * C; c++) {
            double s = 0;
            for (int k = 0; k < 8; k++) s += g_cat2[child(S.D2, p, k) * 6 * C + c];
            g_c2[p * 4 * C + c] = em ? bf16r(s) : s;
        }
    bwd(5, tc1, T_(g_c2, 4 * C), S.D3, &gc1, &tc1);
    bwd(4, T_(S.p2, 2 * C), gc1, S.D3, &gp2, nullptr);
    // pooling backward into the skip2 slice [4C, 6C) of g_cat2, masked by skip2 > 0
    for (int64_t v = 0; v < V2; v++)
        for (int c = 0; c < 2 * C; c++) {
            const double m = S.cat2[v * 6 * C + 4 * C + c];
            double val = m > 0 ? 0.125 * g_p2[parent(S.D2, v) * 2 * C + c] : 0.0;
            val += g_cat2[v * 6 * C + 4 * C + c];
            g_cat2[v * 6 * C + 4 * C + c] = em ? bf16r(val) : val;
        }
    bwd(3, tb1, T_(g_cat2, 6 * C, 4 * C), S.D2, &gb1, &tb1);
    bwd(2, T_(S.p1, C), gb1, S.D2, &gp1, nullptr);
    for (int64_t v = 0; v < V1; v++)
        for (int c = 0; c < C; c++) {
            const double m = S.cat1[v * 3 * C + 2 * C + c];
            double val = m > 0 ? 0.125 * g_p1[parent(S.D1, v) * C + c] : 0.0;
            val += g_cat1[v * 3 * C + 2 * C + c];
            g_cat1[v * 3 * C + 2 * C + c] = em ? bf16r(val) : val;
        }
    bwd(1, ta1, T_(g_cat1, 3 * C, 2 * C), S.D1, &ga1, &ta1);
    bwd(0, T_(S.x8, 8), ga1, S.D1, nullptr, nullptr);
    return loss / (double)V1;
}

}  // extern "C"
