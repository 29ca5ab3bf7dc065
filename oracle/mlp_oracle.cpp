// CPU oracle of the warm-start network (no reference counterpart).
//
// TEST INFRASTRUCTURE ONLY: the GPU forward / backward / optimizer kernels
// are checked against this FP64 statement, and it against finite differences.
//
// Network: 30 -> 64 -> 64 -> 4, ReLU after the two hidden layers, linear
// output. Parameters packed as W1[30][64] b1[64] W2[64][64] b2[64] W3[64][4]
// b3[4] (6404 values). Loss: mean over the batch and the 4 outputs of the
// squared error. Adam with bias correction.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
constexpr int NI = 30, NH = 64, NO = 4;
constexpr int OW1 = 0, OB1 = OW1 + NI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NO, NP = OB3 + NO;

struct Act {
    std::vector<double> z1, h1, z2, h2, y;
};

void forward(const double* w, const double* x, int n, Act& a) {
    a.z1.assign((size_t)n * NH, 0);
    a.h1.assign((size_t)n * NH, 0);
    a.z2.assign((size_t)n * NH, 0);
    a.h2.assign((size_t)n * NH, 0);
    a.y.assign((size_t)n * NO, 0);
    for (int r = 0; r < n; ++r) {
        const double* xr = x + (size_t)r * NI;
        for (int j = 0; j < NH; ++j) {
            double s = w[OB1 + j];
            for (int k = 0; k < NI; ++k) s += xr[k] * w[OW1 + k * NH + j];
            a.z1[r * NH + j] = s;
            a.h1[r * NH + j] = s > 0 ? s : 0;
        }
        for (int j = 0; j < NH; ++j) {
            double s = w[OB2 + j];
            for (int k = 0; k < NH; ++k) s += a.h1[r * NH + k] * w[OW2 + k * NH + j];
            a.z2[r * NH + j] = s;
            a.h2[r * NH + j] = s > 0 ? s : 0;
        }
        for (int j = 0; j < NO; ++j) {
            double s = w[OB3 + j];
            for (int k = 0; k < NH; ++k) s += a.h2[r * NH + k] * w[OW3 + k * NO + j];
            a.y[r * NO + j] = s;
        }
    }
}

double loss_grad(const double* w, const double* x, const double* t, int n, double* g) {
    Act a;
    forward(w, x, n, a);
    std::memset(g, 0, sizeof(double) * NP);
    double loss = 0;
    const double sc = 1.0 / ((double)n * NO);
    std::vector<double> dy(NO), dh2(NH), dz2(NH), dh1(NH), dz1(NH);
    for (int r = 0; r < n; ++r) {
        for (int j = 0; j < NO; ++j) {
            const double e = a.y[r * NO + j] - t[r * NO + j];
            loss += e * e * sc;
            dy[j] = 2.0 * e * sc;
        }
        for (int j = 0; j < NO; ++j) {
            g[OB3 + j] += dy[j];
            for (int k = 0; k < NH; ++k) g[OW3 + k * NO + j] += a.h2[r * NH + k] * dy[j];
        }
        for (int k = 0; k < NH; ++k) {
            double s = 0;
            for (int j = 0; j < NO; ++j) s += w[OW3 + k * NO + j] * dy[j];
            dh2[k] = s;
            dz2[k] = a.z2[r * NH + k] > 0 ? s : 0;
        }
        for (int j = 0; j < NH; ++j) {
            g[OB2 + j] += dz2[j];
            for (int k = 0; k < NH; ++k) g[OW2 + k * NH + j] += a.h1[r * NH + k] * dz2[j];
        }
        for (int k = 0; k < NH; ++k) {
            double s = 0;
            for (int j = 0; j < NH; ++j) s += w[OW2 + k * NH + j] * dz2[j];
            dh1[k] = s;
            dz1[k] = a.z1[r * NH + k] > 0 ? s : 0;
        }
        for (int j = 0; j < NH; ++j) {
            g[OB1 + j] += dz1[j];
            for (int k = 0; k < NI; ++k) g[OW1 + k * NH + j] += x[(size_t)r * NI + k] * dz1[j];
        }
    }
    return loss;
}
}  // namespace

extern "C" {

int32_t orc_mlp_nparams() { return NP; }

// y[n][4] = network(x[n][30])
void orc_mlp_forward(const double* w, const double* x, int32_t n, double* y) {
    Act a;
    forward(w, x, n, a);
    std::memcpy(y, a.y.data(), sizeof(double) * (size_t)n * NO);
}

// loss and its gradient (g[6404]) for targets t[n][4]
double orc_mlp_loss_grad(const double* w, const double* x, const double* t, int32_t n, double* g) {
    return loss_grad(w, x, t, n, g);
}

// one Adam step (step counts from 1): m, v moment buffers updated in place
void orc_mlp_adam(double* w, double* m, double* v, const double* g, int32_t step, double lr, double b1, double b2,
                  double eps) {
    const double c1 = 1.0 - std::pow(b1, step), c2 = 1.0 - std::pow(b2, step);
    for (int i = 0; i < NP; ++i) {
        m[i] = b1 * m[i] + (1 - b1) * g[i];
        v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
        w[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
    }
}

}  // extern "C"
