// bifhip_dev.h -- device-side building blocks shared by the gfx950 kernels.
//
// Everything here is compiled with -ffp-contract=off: an expression written
// as a*b + c is a multiply and an add, each rounded, exactly as in the
// reference's host-compiled semantics.  Fast paths ask for fused
// multiply-adds explicitly with fma()/fmaf().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bifhip {

constexpr int kMaxVal = 16;   // parameter vector capacity
constexpr int kWave = 64;     // CDNA wavefront width

// Model ids (the public numbering, see include/bifhip.h).
enum ModelId : int { kRossler = 0, kLinear = 1, kMemristive = 2 };

// Grid description handed to every map kernel by value.
//   value of swept parameter k for flattened cell idx (getValueByIdx,
//   Library/cudaLibrary.cu:210-216):
//     start[k] + (double)(((int)((float)idx / pk[k])) % nPts) * step[k]
//   pk[k] = powf(nPts, k) and step[k] = (end-start)/(nPts-1) are evaluated
//   on the host with the same IEEE operations the reference uses.
struct Grid {
    int nPts;
    int dimension;
    int icSweep;      // 1: swept entries index the initial state, 0: the parameters
    int mut[2];
    float pk[2];
    double start[2];
    double step[2];
    int nIC;
    int nVal;
    double ic[8];
    double values[kMaxVal];
};

// Arguments of the bifurcation kernels (k_bif.hip).
struct BifArgs {
    Grid g;
    double h;
    double maxValue;
    double hpf;          // peakFinder's h (Library/cudaLibrary.cuh:204)
    int cellBase;        // flattened index of launch cell 0 (amountOfCalculatedPoints)
    const int* cellList; // optional launch-cell -> flattened index map
    int nCells;
    int skip;            // transient steps
    int N;               // recorded samples (amountOfPointsInBlock)
    int pre;             // preScaller
    int cap;             // peak slots per cell row
    long long rowStride; // doubles between cell rows
};

// Arguments of the fused LLE + peak-count kernel (k_bif.hip).
struct FusedArgs {
    BifArgs pk;           // peak side: grid, h, skip (window start), N, cap
    double eps, tMax;
    int skipL, nNT, nBlocks;
    const double* dirs;   // LLE directions per launch cell (null = along x)
};

// Arguments of the Lyapunov kernels (k_lyap.hip).
struct LyapArgs {
    Grid g;
    double h;
    double eps;
    double tMax;
    double maxValue;
    int cellBase;
    const int* cellList;
    int nCells;
    int skip;            // transient steps
    int nNT;             // steps between renormalisations, (size_t)(NT/h)
    int nBlocks;         // renormalisations, (int)(tMax/NT)
    const double* dirs;  // per launch cell: 3 (LLE) or 9 (LS) doubles; null = axis
};

// Arguments of the persistent tile kernel (k_tiles.hip).
// tile kernel variants
constexpr int kTileFp32 = 0;     // FP32 state, one lane per cell
constexpr int kTileFp32x3 = 1;   // FP32 state, three lanes per cell (one tangent vector each)
constexpr int kTileFp64 = 2;     // FP64 state and tangent vectors, one lane per cell

struct TileArgs {
    Grid g;                 // lattice the tile cells live on (nPts x nPts)
    int nTiles;
    int tileW, tileH;
    int itemsPerTile;       // work items per tile
    int itemCells;          // cells per work item = blockDim * cells per lane
    const int* order;       // queue position -> tile id (longest first)
    const int* originI;     // tile origin on the lattice (first swept axis)
    const int* originJ;     // (second swept axis)
    const int* nBlocks;     // renormalisations per tile
    double h;
    int nNT;
    int skip;
    double* out;            // [tile][cell][3]
    int* counter;           // work-queue head, zeroed by the launch
};

// Flattened cell index -> parameter value (float division is IEEE-correct
// under -fhip-fp32-correctly-rounded-divide-sqrt).
__device__ __forceinline__ double grid_value(const Grid& g, int idx, int k)
{
    float q = (float)idx / g.pk[k];
    int t = (int)q;
    return g.start[k] + (double)(t % g.nPts) * g.step[k];
}

// ---------------------------------------------------------------- models
//
// Rossler, Library/cudaLibrary.cu:15-17: sequential explicit Euler.  The
// exact form keeps the reference's operation order (12 roundings/step).
struct Rossler {
    static constexpr int D = 3;
    struct P64 { double a, b, c; };
    struct P32 { float a, b, c, h; float ha1p, hb; };

    __device__ __forceinline__ static P64 params(const double* v) { return {v[0], v[1], v[2]}; }

    __device__ __forceinline__ static void step(double& x0, double& x1, double& x2,
                                                const P64& p, double h)
    {
        x0 = x0 + h * (-x1 - x2);
        x1 = x1 + h * (x0 + p.a * x1);
        x2 = x2 + h * (p.b + x2 * (x0 - p.c));
    }

    // Two independent trajectories (x with parameters p, y with q) advanced
    // one step each, written operation by operation in alternation so the
    // two dependency chains interleave in the instruction stream; every
    // operation and its operand order are those of step(), so each result
    // is bit-identical to step().
    __device__ __forceinline__ static void step2(double& x0, double& x1, double& x2,
                                                 const P64& p, double& y0, double& y1,
                                                 double& y2, const P64& q, double h)
    {
        double ta = -x1 - x2, tb = -y1 - y2;
        ta = h * ta;
        tb = h * tb;
        x0 = x0 + ta;
        y0 = y0 + tb;
        double ua = p.a * x1, ub = q.a * y1;
        ua = x0 + ua;
        ub = y0 + ub;
        ua = h * ua;
        ub = h * ub;
        x1 = x1 + ua;
        y1 = y1 + ub;
        double va = x0 - p.c, vb = y0 - q.c;
        va = x2 * va;
        vb = y2 * vb;
        va = p.b + va;
        vb = q.b + vb;
        va = h * va;
        vb = h * vb;
        x2 = x2 + va;
        y2 = y2 + vb;
    }

    __device__ __forceinline__ static void stepn(double (&x)[3], const P64& p, double h)
    {
        step(x[0], x[1], x[2], p, h);
    }

    // FP32 fast form of the same map with fused multiply-adds (7 ops):
    //   x0 <- x0 - h(x1+x2)
    //   x1 <- x1(1+ha) + h x0
    //   x2 <- x2(1 + h(x0-c)) + hb
    __device__ __forceinline__ static P32 params32(const double* v, double h)
    {
        P32 q;
        q.a = (float)v[0]; q.b = (float)v[1]; q.c = (float)v[2]; q.h = (float)h;
        q.ha1p = (float)(1.0 + h * v[0]);
        q.hb = (float)(h * v[1]);
        return q;
    }
    __device__ __forceinline__ static void step32(float& x0, float& x1, float& x2, const P32& p)
    {
        x0 = fmaf(-p.h, x1 + x2, x0);
        x1 = fmaf(x1, p.ha1p, p.h * x0);
        x2 = fmaf(x2, fmaf(p.h, x0 - p.c, 1.0f), p.hb);
    }
    // Deviation u (true deviation = s*u) of the FP32 map along (x -> x'),
    // exact difference of two trajectories x and x + s*u, given the
    // pre-step x2 (x2o) and post-step x0 (x0n):
    //   u0' = u0 - h(u1+u2)
    //   u1' = u1(1+ha) + h u0'
    //   u2' = u2 + h(u2 (x0n + s u0' - c) + x2o u0')
    __device__ __forceinline__ static void dev32(float& u0, float& u1, float& u2, float x0n,
                                                 float x2o, float s, const P32& p)
    {
        u0 = fmaf(-p.h, u1 + u2, u0);
        u1 = fmaf(u1, p.ha1p, p.h * u0);
        float g = fmaf(s, u0, x0n - p.c);
        u2 = fmaf(p.h, fmaf(u2, g, x2o * u0), u2);
    }
    // Tangent map (variational equation of the discrete step), FP32.
    //   v0' = v0 - h(v1+v2); v1' = v1(1+ha) + h v0'; v2' = v2(1+h(x0n-c)) + h x2o v0'
    __device__ __forceinline__ static void tan32(float& v0, float& v1, float& v2, float x0n,
                                                 float x2o, const P32& p)
    {
        v0 = fmaf(-p.h, v1 + v2, v0);
        v1 = fmaf(v1, p.ha1p, p.h * v0);
        v2 = fmaf(v2, fmaf(p.h, x0n - p.c, 1.0f), (p.h * x2o) * v0);
    }
    // Tangent map of `step` in FP64, one rounding per operation in this
    // order (no contraction): given the post-step x0 (x0n) and the pre-step
    // x2 (x2o),
    //   v0' = v0 - h(v1+v2); v1' = (1+ha)v1 + h v0'; v2' = (1+h(x0n-c))v2 + h x2o v0'
    __device__ __forceinline__ static void tan64(double& v0, double& v1, double& v2, double x0n,
                                                 double x2o, const P64& p, double h)
    {
        v0 = v0 - h * (v1 + v2);
        v1 = (1 + h * p.a) * v1 + h * v0;
        v2 = (1 + h * (x0n - p.c)) * v2 + h * x2o * v0;
    }
};

// Linear diagonal test model x_i <- x_i + h*(a_i*x_i); its Lyapunov
// spectrum is log|1+h a_i|/h.  Used only by the closed-form tests.
struct Linear {
    static constexpr int D = 3;
    struct P64 { double a, b, c; };
    struct P32 { float a, b, c, h; float ha1p, hb; };
    __device__ __forceinline__ static P64 params(const double* v) { return {v[0], v[1], v[2]}; }
    __device__ __forceinline__ static void step(double& x0, double& x1, double& x2,
                                                const P64& p, double h)
    {
        x0 = x0 + h * (p.a * x0);
        x1 = x1 + h * (p.b * x1);
        x2 = x2 + h * (p.c * x2);
    }
    __device__ __forceinline__ static void step2(double& x0, double& x1, double& x2,
                                                 const P64& p, double& y0, double& y1,
                                                 double& y2, const P64& q, double h)
    {
        step(x0, x1, x2, p, h);
        step(y0, y1, y2, q, h);
    }
    __device__ __forceinline__ static void stepn(double (&x)[3], const P64& p, double h)
    {
        step(x[0], x[1], x[2], p, h);
    }
    __device__ __forceinline__ static P32 params32(const double* v, double h)
    {
        P32 q;
        q.a = (float)(1.0 + h * v[0]); q.b = (float)(1.0 + h * v[1]);
        q.c = (float)(1.0 + h * v[2]); q.h = (float)h; q.ha1p = 0; q.hb = 0;
        return q;
    }
    __device__ __forceinline__ static void step32(float& x0, float& x1, float& x2, const P32& p)
    {
        x0 *= p.a; x1 *= p.b; x2 *= p.c;
    }
    __device__ __forceinline__ static void dev32(float& u0, float& u1, float& u2, float, float,
                                                 float, const P32& p)
    {
        u0 *= p.a; u1 *= p.b; u2 *= p.c;
    }
    __device__ __forceinline__ static void tan32(float& v0, float& v1, float& v2, float, float,
                                                 const P32& p)
    {
        v0 *= p.a; v1 *= p.b; v2 *= p.c;
    }
    __device__ __forceinline__ static void tan64(double& v0, double& v1, double& v2, double,
                                                 double, const P64& p, double h)
    {
        step(v0, v1, v2, p, h);
    }
};

// Five-variable memristive system, the active body of the reference's
// calculateDiscreteModel (Library/cudaLibrary.cu:19-43): a composition
// step of an explicit sequential pass of length k = h*a7 (k held in a
// float, Library/cudaLibrary.cu:19-27), the state then rounded to float
// (y[5], :29-33), and a semi-implicit pass of length h*(1-a7) whose linear
// terms are solved in closed form (:36-43).  Parameters a0..a7; |x4|
// (double absolute value) enters the dissipation term.  Every operation in
// the reference's order, no contraction.
struct Memristive {
    static constexpr int D = 5;
    struct P64 { double a[8]; };

    __device__ __forceinline__ static P64 params(const double* v)
    {
        P64 p;
#pragma unroll
        for (int i = 0; i < 8; ++i) p.a[i] = v[i];
        return p;
    }

    __device__ __forceinline__ static void stepn(double (&x)[5], const P64& p, double h)
    {
        const double* a = p.a;
        const float k = (float)(h * a[7]);
        x[4] = x[4] + k * (x[0]);
        x[3] = x[3] + k * (-a[2] * x[1] - a[3] * x[3]);
        x[2] = x[2] + k * (a[1] * (x[1] - x[0] - x[2]));
        x[1] = x[1] + k * (-x[2] + x[3]);
        x[0] = x[0] + k * (a[0] * (x[2] - (a[4] + a[5] + a[6] * fabs(x[4])) * x[0]));
        const float y0 = (float)x[0], y1 = (float)x[1], y2 = (float)x[2], y3 = (float)x[3],
                    y4 = (float)x[4];
        const double h1 = h * (1 - a[7]);
        x[0] = (y0 + h1 * (a[0] * x[2])) / (1 + h1 * a[0] * (a[4] + a[5] + a[6] * fabs(x[4])));
        x[1] = y1 + h1 * (-x[2] + x[3]);
        x[2] = (y2 + h1 * (a[1] * (x[1] - x[0]))) / (1 + h1 * a[1]);
        x[3] = (y3 + h1 * (-a[2] * x[1])) / (1 + h1 * a[3]);
        x[4] = y4 + h1 * (x[0]);
    }
};

// Parameter vector and initial state (D variables) of a cell, swept
// entries substituted as in cell_setup.
template <int D>
__device__ __forceinline__ void cell_setup_n(const Grid& g, int idx, double* v, double (&x)[D])
{
#pragma unroll
    for (int i = 0; i < kMaxVal; ++i) v[i] = g.values[i];
#pragma unroll
    for (int i = 0; i < D; ++i) x[i] = g.ic[i];
    for (int k = 0; k < g.dimension; ++k) {
        const double val = grid_value(g, idx, k);
        if (g.icSweep) {
#pragma unroll
            for (int i = 0; i < D; ++i)
                if (i == g.mut[k]) x[i] = val;
        } else {
#pragma unroll
            for (int i = 0; i < kMaxVal; ++i)
                if (i == g.mut[k]) v[i] = val;
        }
    }
}

// Parameter vector and initial state of a cell, the swept entries
// substituted into the parameters (calculateDiscreteModelCUDA,
// Library/cudaLibrary.cu:96-105) or, for a sweep over initial conditions,
// into the state (calculateDiscreteModelICCUDA / LLEKernelICCUDA,
// Library/cudaLibrary.cu:190-192, 837-839).
__device__ __forceinline__ void cell_setup(const Grid& g, int idx, double* v, double& x0,
                                           double& x1, double& x2)
{
#pragma unroll
    for (int i = 0; i < kMaxVal; ++i) v[i] = g.values[i];
    x0 = g.ic[0];
    x1 = g.ic[1];
    x2 = g.ic[2];
    for (int k = 0; k < g.dimension; ++k) {
        const double val = grid_value(g, idx, k);
        // dynamic index into a small register array: keep it a select chain
        if (g.icSweep) {
            if (g.mut[k] == 0) x0 = val;
            else if (g.mut[k] == 1) x1 = val;
            else if (g.mut[k] == 2) x2 = val;
        } else {
#pragma unroll
            for (int i = 0; i < kMaxVal; ++i)
                if (i == g.mut[k]) v[i] = val;
        }
    }
}

// Divergence test of loopCalculateDiscreteModel (Library/cudaLibrary.cu:69-71):
// |(float)x| > maxValue in double, only when maxValue != 0.
__device__ __forceinline__ bool diverged(double xw, double maxValue)
{
    return (double)fabsf((float)xw) > maxValue;
}

// Sum of log(norm) over many renormalisations, kept as the product of the
// squared norms in FP64 (mantissa and integer exponent) and logged once.  A
// product of N norms rounds N times: about N * 1.1e-16 in the log sum
// (4.4e-13 for 4 000 renormalisations).  A zero norm gives -inf, NaN and
// inf propagate, as a sum of logs does.
struct LogNormSum {
    double m;
    int e;
    __device__ __forceinline__ void init()
    {
        m = 1.0;
        e = 0;
    }
    __device__ __forceinline__ void add(double d2)
    {
        int ea, eb;
        const double a = frexp(d2, &ea);
        m = frexp(m * a, &eb);
        e += ea + eb;
    }
    __device__ __forceinline__ double value() const
    {
        return 0.5 * (log(m) + (double)e * 0.69314718055994530942);
    }
};

// Classical Gram-Schmidt of three row vectors a -> b with norms
// (projectionOperator/gramSchmidtProcess, Library/cudaLibrary.cu:913-957).
__device__ __forceinline__ void gs3(const double (&a)[9], double (&b)[9], double (&den)[3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) b[i * 3 + j] = a[i * 3 + j];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            double num = 0, dd = 0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                num += a[i * 3 + m] * b[j * 3 + m];
                dd += b[j * 3 + m] * b[j * 3 + m];
            }
            const double f = dd == 0 ? 0 : num / dd;
#pragma unroll
            for (int m = 0; m < 3; ++m) b[i * 3 + m] -= f * b[j * 3 + m];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double d = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) d += b[i * 3 + j] * b[i * 3 + j];
        d = sqrt(d);
#pragma unroll
        for (int j = 0; j < 3; ++j) b[i * 3 + j] = d == 0 ? 0 : b[i * 3 + j] / d;
        den[i] = d;
    }
}

// ------------------------------------------------------- wave utilities

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

}  // namespace bifhip
