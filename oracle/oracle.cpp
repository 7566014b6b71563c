// oracle.cpp -- CPU FP64 restatement of pimc's all-bead-pairs hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker for libpimc_gpu.so.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
//
// Single-threaded FP64, built with -ffp-contract=off so every expression is
// evaluated in the order it is written, which is the order of the cited
// reference source.  All citations are relative to the reference repository.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr double kEPS = 1.0e-7;      // include/common.h:89
constexpr int kXXX = -1;             // include/common.h:93
constexpr int kNPCFSEP = 50;         // include/common.h:85
constexpr int kNDIM = 3;

using Vec = std::array<double, 3>;
using IVec = std::array<int, 3>;
using Bead = std::array<int, 2>;

inline double dot3(const Vec& a, const Vec& b) {
    // include/array_math.h:210-216: sequential accumulation from 0
    double s = 0.0;
    s += a[0] * b[0];
    s += a[1] * b[1];
    s += a[2] * b[2];
    return s;
}

// ---------------------------------------------------------------------------
// MT19937 exactly as MTRand (include/MersenneTwister.h).  The reference holds
// the 624-word state in `unsigned long` masked to 32 bits; all of its updates
// are congruences mod 2^32, so uint32_t arithmetic gives the same words.
// ---------------------------------------------------------------------------
struct MT {
    static constexpr int N = 624, M = 397;
    uint32_t s[N];
    int left = 0;
    int next = 0;

    explicit MT(uint32_t seed) {
        // initialize (:173-188)
        s[0] = seed;
        for (int i = 1; i < N; ++i)
            s[i] = 1812433253u * (s[i - 1] ^ (s[i - 1] >> 30)) + static_cast<uint32_t>(i);
        reload();
    }
    static uint32_t twist(uint32_t m, uint32_t a, uint32_t b) {
        uint32_t mix = (a & 0x80000000u) | (b & 0x7fffffffu);
        return m ^ (mix >> 1) ^ ((b & 1u) ? 0x9908b0dfu : 0u);
    }
    void reload() {
        // reload (:190-204)
        int i = 0;
        for (; i < N - M; ++i) s[i] = twist(s[i + M], s[i], s[i + 1]);
        for (; i < N - 1; ++i) s[i] = twist(s[i + M - N], s[i], s[i + 1]);
        s[N - 1] = twist(s[M - 1], s[N - 1], s[0]);
        left = N;
        next = 0;
    }
    uint32_t randInt() {
        // (:290-304)
        if (left == 0) reload();
        --left;
        uint32_t y = s[next++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        return y ^ (y >> 18);
    }
    uint32_t randInt(uint32_t n) {
        // masked rejection (:306-323)
        uint32_t used = n;
        used |= used >> 1;
        used |= used >> 2;
        used |= used >> 4;
        used |= used >> 8;
        used |= used >> 16;
        uint32_t i;
        do { i = randInt() & used; } while (i > n);
        return i;
    }
    double rand() { return double(randInt()) * (1.0 / 4294967295.0); }
    double randExc() { return double(randInt()) * (1.0 / 4294967296.0); }
    double randNorm(double mean, double sd) {
        // Marsaglia polar form (:349-363)
        double x, y, r;
        do {
            x = 2.0 * rand() - 1.0;
            y = 2.0 * rand() - 1.0;
            r = x * x + y * y;
        } while (r >= 1.0 || r == 0.0);
        double f = std::sqrt(-2.0 * std::log(r) / r);
        return mean + x * f * sd;
    }
};

// ---------------------------------------------------------------------------
// Prism container (src/container.cpp:84-109, include/container.h:50-59)
// ---------------------------------------------------------------------------
struct Box {
    Vec side{}, sideInv{}, pSide{};
    double volume = 0, maxSep = 0;
    void init(double L) {
        for (int i = 0; i < 3; ++i) {
            side[i] = L;
            sideInv[i] = 1.0 / L;
            pSide[i] = L;
        }
        Vec h{side[0] / 2.0, side[1] / 2.0, side[2] / 2.0};
        maxSep = std::sqrt(dot3(h, h));
        volume = side[0] * side[1] * side[2];
    }
    void putInBC(Vec& r) const {
        for (int i = 0; i < 3; ++i) r[i] -= pSide[i] * std::floor(r[i] * sideInv[i] + 0.5);
    }
};

// ---------------------------------------------------------------------------
// Aziz HFDHE2 (src/potential.cpp:1741-1909) + TabulatedPotential direct lookup
// (include/potential.h:163-260).
// ---------------------------------------------------------------------------
struct Aziz {
    int year = 1979;
    double rm = 0, A = 0, epsilon = 0, alpha = 0, beta = 0, D = 0, C6 = 0, C8 = 0, C10 = 0;
    double dr = 0;
    int tableLength = 0;
    std::vector<double> tV, tdV, td2V;
    double tailV = 0;

    void setYear(int year) {
        if (year == 1987) {
            epsilon = 10.948; rm = 2.9673; D = 1.4826; alpha = 10.43329537; beta = -2.27965105;
            C6 = 1.36745214; C8 = 0.42123807; C10 = 0.17473318; A = 1.8443101E5;
        } else if (year == 1995) {
            epsilon = 10.956; rm = 2.9683; D = 1.438; alpha = 10.5717543; beta = -2.07758779;
            C6 = 1.35186623; C8 = 0.4149514; C10 = 0.17151143; A = 1.86924404E5;
        } else {  // 1979
            epsilon = 10.8; rm = 2.9673; D = 1.241314; alpha = 13.353384; beta = 0.0;
            C6 = 1.3732412; C8 = 0.4253785; C10 = 0.1781; A = 0.5448504E6;
        }
    }
    double F(double x) const { return (x < D ? std::exp(-(D / x - 1.0) * (D / x - 1.0)) : 1.0); }
    double dF(double x) const {
        double ix = 1.0 / x;
        double r = 2.0 * D * ix * ix * (D * ix - 1.0) * std::exp(-(D * ix - 1.0) * (D * ix - 1.0));
        return (x < D ? r : 0.0);
    }
    double d2F(double x) const {
        double ix = 1.0 / x;
        double r = 2.0 * D * ix * ix * ix * (2.0 * D * D * D * ix * ix * ix - 4.0 * D * D * ix * ix - D * ix + 2.0) *
                   std::exp(-(D * ix - 1.0) * (D * ix - 1.0));
        return (x < D ? r : 0.0);
    }
    double valueV(double r) const {
        double x = r / rm;
        double Urep = A * std::exp(-alpha * x + beta * x * x);
        if (x < kEPS) return 0.0;
        if (x < 0.01) return (epsilon * Urep);
        double ix2 = 1.0 / (x * x);
        double ix6 = ix2 * ix2 * ix2;
        double ix8 = ix6 * ix2;
        double ix10 = ix8 * ix2;
        double Uatt = -(C6 * ix6 + C8 * ix8 + C10 * ix10) * F(x);
        return (epsilon * (Urep + Uatt));
    }
    double valuedVdr(double r) const {
        double x = r / rm;
        double T1 = A * (-alpha + 2.0 * beta * x) * std::exp(-alpha * x + beta * x * x);
        if (x < kEPS) return 0.0;
        if (x < 0.01) return ((epsilon / rm) * T1);
        double ix = 1.0 / x;
        double ix2 = ix * ix;
        double ix6 = ix2 * ix2 * ix2;
        double ix7 = ix6 * ix;
        double ix8 = ix6 * ix2;
        double ix9 = ix8 * ix;
        double ix10 = ix8 * ix2;
        double ix11 = ix10 * ix;
        double T2 = (6.0 * C6 * ix7 + 8.0 * C8 * ix9 + 10.0 * C10 * ix11) * F(x);
        double T3 = -(C6 * ix6 + C8 * ix8 + C10 * ix10) * dF(x);
        return ((epsilon / rm) * (T1 + T2 + T3));
    }
    double valued2Vdr2(double r) const {
        double x = r / rm;
        double ab = (alpha - 2.0 * beta * x) * (alpha - 2.0 * beta * x);
        double T1 = A * (2 * beta + ab) * std::exp(-alpha * x + beta * x * x);
        if (x < kEPS) return 0.0;
        if (x < 0.01) return ((epsilon / rm) * T1);
        double ix = 1.0 / x;
        double ix2 = ix * ix;
        double ix6 = ix2 * ix2 * ix2;
        double ix7 = ix6 * ix;
        double ix8 = ix6 * ix2;
        double ix9 = ix8 * ix;
        double ix10 = ix8 * ix2;
        double ix11 = ix10 * ix;
        double ix12 = ix11 * ix;
        double T2 = -(42.0 * C6 * ix8 + 72.0 * C8 * ix10 + 110.0 * C10 * ix12) * F(x);
        double T3 = 2.0 * (6.0 * C6 * ix7 + 8.0 * C8 * ix9 + 10.0 * C10 * ix11) * dF(x);
        double T4 = -(C6 * ix6 + C8 * ix8 + C10 * ix10) * d2F(x);
        return ((epsilon / (rm * rm)) * (T1 + T2 + T3 + T4));
    }
    void init(int year_, double maxSep, double rc, bool derivs) {
        year = year_;
        setYear(year);
        // initLookupTable((1.0E-6)*rm, maxSep): accumulated r += dr (potential.h:163-183)
        dr = (1.0E-6) * rm;
        tableLength = int(maxSep / dr);
        tV.assign(tableLength, 0.0);
        if (derivs) {
            tdV.assign(tableLength, 0.0);
            td2V.assign(tableLength, 0.0);
        }
        double r = 0;
        for (int n = 0; n < tableLength; n++) {
            tV[n] = valueV(r);
            if (derivs) {
                tdV[n] = valuedVdr(r);
                td2V[n] = valued2Vdr2(r);
            }
            r += dr;
        }
        // tail correction (:1800-1806)
        double rmorc = rm / rc;
        double t2 = C6 * std::pow(rmorc, 3.0) / 3.0;
        double t3 = C8 * std::pow(rmorc, 5.0) / 5.0;
        double t4 = C10 * std::pow(rmorc, 7.0) / 7.0;
        tailV = 2.0 * M_PI * epsilon * (-rm * rm * rm * (t2 + t3 + t4));
    }
    double direct(const std::vector<double>& t, double r) const {
        // include/potential.h:249-260; extremal values are 0 for Aziz
        int k = int(r / dr);
        if (k <= 0) return 0.0;
        if (k >= tableLength) return 0.0;
        return t[k];
    }
    double V(const Vec& sep) const { return direct(tV, std::sqrt(dot3(sep, sep))); }
    Vec gradV(const Vec& sep) const {
        double rn = std::sqrt(dot3(sep, sep));
        double f = direct(tdV, rn) / rn;
        return Vec{f * sep[0], f * sep[1], f * sep[2]};
    }
    double grad2V(const Vec& sep) const { return direct(td2V, std::sqrt(dot3(sep, sep))); }
};

enum BeadState { HEADTAIL = 0, SPECIAL = 1, NONE = 2 };

struct State;

// ---------------------------------------------------------------------------
// LookupTable (src/lookuptable.cpp:26-516, include/lookuptable.h)
// ---------------------------------------------------------------------------
struct Lookup {
    const Box* box = nullptr;
    int M = 0, nalloc = 0;
    IVec nGrid{};
    Vec sizeGrid{};
    int cells = 0;
    static constexpr int numNN = 27;
    std::vector<IVec> gridNN;      // [cells][27]
    int cap = 0;                   // hashSize[NDIM+1]
    std::vector<int> hash;         // [cells][M][cap]
    std::vector<int> numLabels;    // [cells][M]
    std::vector<IVec> grid;        // [M][nalloc]
    std::vector<int> beadLabel;    // [M][nalloc]
    double rc2 = 0;
    // interaction list output (beadList/beadSep/numBeads)
    std::vector<Bead> beadList;
    std::vector<Vec> beadSep;
    int numBeads = 0;

    int cellNumber(const IVec& g) const { return (g[0] * nGrid[1] + g[1]) * nGrid[2] + g[2]; }
    IVec cellIndex(int n) const {
        IVec g;
        for (int i = 0; i < 3; i++) {
            int scale = 1;
            for (int j = i + 1; j < 3; j++) scale *= nGrid[j];
            g[i] = (n / scale) % nGrid[i];
        }
        return g;
    }
    IVec gridIndex(const Vec& pos) const {
        // include/lookuptable.h:139-147 (abs of a double: fabs)
        IVec g;
        for (int i = 0; i < 3; i++)
            g[i] = static_cast<int>(std::fabs(pos[i] + 0.5 * box->side[i] - kEPS) / (sizeGrid[i] + kEPS));
        return g;
    }
    int& H(int cell, int slice, int label) { return hash[(size_t(cell) * M + slice) * cap + label]; }
    int& NL(int cell, int slice) { return numLabels[size_t(cell) * M + slice]; }

    void setup(const Box* b, int M_, int nalloc_, double rc) {
        box = b;
        M = M_;
        nalloc = nalloc_;
        // setupNNGrid (:67-100)
        cells = 1;
        for (int i = 0; i < 3; i++) {
            nGrid[i] = static_cast<int>(std::floor((box->side[i] / rc) + kEPS));
            if (nGrid[i] < 1) nGrid[i] = 1;
            sizeGrid[i] = box->side[i] / (1.0 * nGrid[i]);
            cells *= nGrid[i];
        }
        numLabels.assign(size_t(cells) * M, 0);
        // shift vectors (:115-126)
        std::array<IVec, numNN> nnShift{};
        const int sh[3] = {-1, 0, 1};
        for (int n = 0; n < numNN; n++)
            for (int i = 0; i < 3; i++) {
                int index = ((int)std::floor(1.0 * n / std::pow(3.0, 1.0 * (kNDIM - 1 - i)) + kEPS)) % 3;
                nnShift[n][i] = sh[index];
            }
        gridNN.assign(size_t(cells) * numNN, IVec{0, 0, 0});
        for (int n = 0; n < cells; n++) {
            IVec g = cellIndex(n);
            for (int m = 0; m < numNN; m++) {
                IVec t;
                for (int i = 0; i < 3; i++) {
                    t[i] = g[i] + nnShift[m][i];
                    if (t[i] == nGrid[i]) t[i] = 0;                 // periodic - 1
                    else if (t[i] == -1) t[i] = nGrid[i] - 1;      // periodic*n - 1
                }
                gridNN[size_t(n) * numNN + m] = t;
            }
        }
        // duplicates -> -1, keeping the last occurrence (:161-184)
        const IVec neg{-1, -1, -1};
        for (int n = 0; n < cells; n++) {
            IVec* row = &gridNN[size_t(n) * numNN];
            for (int m = numNN - 1; m >= 0; m--) {
                IVec dup = row[m];
                for (int p = m - 1; p >= 0; p--)
                    if (row[p] == dup) row[p] = neg;
            }
        }
        // hash sizing (:34-46)
        cap = static_cast<int>(nalloc / cells) + 5;
        hash.assign(size_t(cells) * M * cap, kXXX);
        grid.assign(size_t(M) * nalloc, IVec{kXXX, kXXX, kXXX});
        beadLabel.assign(size_t(M) * nalloc, kXXX);
        beadList.assign(nalloc, Bead{kXXX, kXXX});
        beadSep.assign(nalloc, Vec{0, 0, 0});
        rc2 = rc * rc;
    }
    void grow() {
        // resizeAndPreserve of the last hash extent by one
        int ncap = cap + 1;
        std::vector<int> nh(size_t(cells) * M * ncap, kXXX);
        for (size_t cs = 0; cs < size_t(cells) * M; cs++)
            for (int l = 0; l < cap; l++) nh[cs * ncap + l] = hash[cs * cap + l];
        hash.swap(nh);
        cap = ncap;
    }
    void append(const Bead& b) {
        int cell = cellNumber(grid[size_t(b[0]) * nalloc + b[1]]);
        int label = NL(cell, b[0]);
        if (cap == label) grow();
        H(cell, b[0], label) = b[1];
        beadLabel[size_t(b[0]) * nalloc + b[1]] = label;
        NL(cell, b[0])++;
    }
    void delBead(const Bead& b) {
        // (:430-458)
        size_t bi = size_t(b[0]) * nalloc + b[1];
        int label = beadLabel[bi];
        int cell = cellNumber(grid[bi]);
        int lastLabel = NL(cell, b[0]) - 1;
        if (label < lastLabel) {
            int moved = H(cell, b[0], lastLabel);
            H(cell, b[0], lastLabel) = kXXX;
            H(cell, b[0], label) = moved;
            beadLabel[size_t(b[0]) * nalloc + moved] = label;
        } else {
            H(cell, b[0], lastLabel) = kXXX;
        }
        beadLabel[bi] = kXXX;
        grid[bi] = IVec{kXXX, kXXX, kXXX};
        NL(cell, b[0])--;
    }
    void updateBead(const Bead& b, const Vec& pos) {
        // (:355-393)
        IVec g = gridIndex(pos);
        size_t bi = size_t(b[0]) * nalloc + b[1];
        if (!(g == grid[bi])) {
            delBead(b);
            grid[bi] = g;
            append(b);
        }
    }
    void updateGrid(const State& st);
    void updateInteractionList(const State& st, const Bead& b1);
};

// ---------------------------------------------------------------------------
// Path + Worm + LocalAction state
// ---------------------------------------------------------------------------
struct State {
    Box box;
    int M = 0, nalloc = 0;
    double tau = 0, lambda = 0, rc = 0;
    std::vector<Vec> beads;        // [M][nalloc]
    std::vector<uint8_t> on;       // worm bead mask
    std::vector<int> numBeadsAtSlice;
    std::vector<Bead> nextL, prevL;
    Bead head{kXXX, kXXX}, tail{kXXX, kXXX}, sp1{kXXX, kXXX}, sp2{kXXX, kXXX};
    Aziz aziz;
    Lookup lookup;
    int shift = 1;
    double VFactor[2] = {1.0, 1.0};
    double gradVFactor[2] = {0.0, 0.0};   // GSF / Li-Broughton correction (setup.cpp:1232-1256)
    int period = 1;
    std::array<int, kNPCFSEP> sepHist{};
    double dSep = 0;

    size_t idx(const Bead& b) const { return size_t(b[0]) * nalloc + b[1]; }
    const Vec& pos(const Bead& b) const { return beads[idx(b)]; }
    bool beadOn(const Bead& b) const { return on[idx(b)] != 0; }
    const Bead& next(const Bead& b) const { return nextL[idx(b)]; }
    const Bead& prev(const Bead& b) const { return prevL[idx(b)]; }
    Bead next(const Bead& b, int n) const {
        Bead r = b;
        for (int m = 0; m < n; m++) r = next(r);
        return r;
    }
    Bead prev(const Bead& b, int n) const {
        Bead r = b;
        for (int m = 0; m < n; m++) r = prev(r);
        return r;
    }
    int numBeadsOn() const {
        int s = 0;
        for (auto v : on) s += v;
        return s;
    }
    int trueNumParticles() const { return numBeadsOn() / M; }
    Vec getSeparation(const Bead& b1, const Bead& b2) const {
        // include/path.h:179-184
        const Vec& p1 = pos(b1);
        const Vec& p2 = pos(b2);
        Vec s{p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
        box.putInBC(s);
        return s;
    }
    Vec getVelocity(const Bead& b) const {
        // include/path.h:189-203
        if ((b[0] == kXXX && b[1] == kXXX)) return Vec{0, 0, 0};
        const Bead& nb = next(b);
        if (nb[0] == kXXX && nb[1] == kXXX) return Vec{0, 0, 0};
        const Vec& p1 = pos(nb);
        const Vec& p0 = pos(b);
        Vec v{p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        box.putInBC(v);
        return v;
    }
    BeadState getState(const Bead& b) const {
        // include/worm.h:110-117
        if (b == head || b == tail) return HEADTAIL;
        if (b == sp1 || b == sp2) return SPECIAL;
        return NONE;
    }
    double factor(BeadState s1) const { return 1.0 - 0.5 * (s1 != NONE); }
    double factor(BeadState s1, const Bead& b2) const {
        // src/worm.cpp:108-118
        BeadState s2 = getState(b2);
        if (s1 == NONE && s2 == NONE) return 1.0;
        if (s1 == HEADTAIL && s2 == HEADTAIL) return 0.0;
        return 0.5;
    }
    double tauShift() const { return shift * tau; }  // include/action.h:136

    void updateSepHist(const Vec& s) {
        // action.cpp:216-224
        int nR = int(std::sqrt(dot3(s, s)) / dSep);
        if (nR >= 0 && nR < kNPCFSEP) ++sepHist[nR];
    }
    std::array<double, 2> Vslice(int slice) {
        // action.cpp:902-947 (the free external potential is identically 0)
        sepHist.fill(0);
        return VslicePairs(slice, true);
    }
    // The pair sum of V(slice); the separation histogram is filled when hist is
    // set (the const form, hist = false, may run on several threads at once).
    std::array<double, 2> VslicePairs(int slice, bool hist) {
        double totVint = 0.0;
        double totVext = 0.0;
        Bead b1{slice, 0}, b2{slice, 0};
        int n = numBeadsAtSlice[slice];
        std::vector<double> extFactors;
        for (b1[1] = 0; b1[1] < n; b1[1]++) {
            BeadState s1 = getState(b1);
            extFactors.push_back(factor(s1));
            for (b2[1] = b1[1] + 1; b2[1] < n; b2[1]++) {
                Vec sep = getSeparation(b2, b1);
                if (hist) updateSepHist(sep);
                totVint += factor(s1, b2) * aziz.V(sep);
            }
        }
        for (double f : extFactors) totVext += f * 0.0;
        return {totVext, totVint};
    }
    // totalAction with the slices' pair sums spread over `threads` host threads
    // (slice m on thread m mod threads) and added in slice order as totalAction
    // adds them: the same value.  Primitive factors only (no gradient correction).
    double totalActionThreads(int threads) {
        std::vector<std::array<double, 2>> v(M);
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t)
            pool.emplace_back([this, t, threads, &v] {
                for (int slice = t; slice < M; slice += threads) v[slice] = VslicePairs(slice, false);
            });
        for (auto& th : pool) th.join();
        double totU = 0.0;
        for (int slice = 0; slice < M; slice++) {
            double s = 0.0;
            s += v[slice][0];
            s += v[slice][1];
            totU += VFactor[slice % 2] * tauShift() * s;
        }
        return totU;
    }
    double Vnn(const Bead& b1, double totVext) {
        // action.cpp:1044-1060
        if (!beadOn(b1)) return 0.0;
        lookup.updateInteractionList(*this, b1);
        BeadState s1 = getState(b1);
        double totVint = 0.0;
        for (int n = 0; n < lookup.numBeads; n++)
            totVint += factor(s1, lookup.beadList[n]) * aziz.V(lookup.beadSep[n]);
        return (totVext + totVint);
    }
    double Vnn(const Bead& b1) {
        // action.cpp:1026-1042
        if (!beadOn(b1)) return 0.0;
        BeadState s1 = getState(b1);
        double totVext = factor(s1) * 0.0;
        return Vnn(b1, totVext);
    }
    double VnnSlice(int slice) {
        // action.cpp:1069-1123
        int n = numBeadsAtSlice[slice];
        std::vector<char> doP(n, 1);
        double totVint = 0.0, totVext = 0.0;
        Bead b1{slice, 0};
        std::vector<double> extFactors;
        for (b1[1] = 0; b1[1] < n; b1[1]++) {
            doP[b1[1]] = 0;
            BeadState s1 = getState(b1);
            extFactors.push_back(factor(s1));
            lookup.updateInteractionList(*this, b1);
            for (int k = 0; k < lookup.numBeads; k++) {
                Bead b2 = lookup.beadList[k];
                if (doP[b2[1]]) {
                    Vec sep = getSeparation(b2, b1);
                    totVint += factor(s1, b2) * aziz.V(sep);
                }
            }
        }
        for (double f : extFactors) totVext += f * 0.0;
        return (totVext + totVint);
    }
    double gradVSquared(int slice) const {
        // action.cpp:1188-1223: sum over the slice's beads of |sum_j gradV(r1 - r2)|^2
        // (the free external potential has gradV = 0)
        double totF2 = 0.0;
        int n = numBeadsAtSlice[slice];
        Bead b1{slice, 0}, b2{slice, 0};
        for (b1[1] = 0; b1[1] < n; b1[1]++) {
            Vec F{0.0, 0.0, 0.0};
            for (b2[1] = 0; b2[1] < n; b2[1]++) {
                if (!(b1 == b2)) {
                    Vec g = aziz.gradV(getSeparation(b1, b2));
                    for (int i = 0; i < 3; i++) F[i] += g[i];
                }
            }
            Vec Fext{0.0, 0.0, 0.0};
            for (int i = 0; i < 3; i++) F[i] += Fext[i];
            totF2 += dot3(F, F);
        }
        return totF2;
    }
    double gradVnnSquared(const Bead& b1) const {
        // action.cpp:1282-1337, over the interaction list left by the preceding Vnn(b1)
        double totF2 = 0.0;
        if (beadOn(b1)) {
            const Vec Fext1{0.0, 0.0, 0.0};
            Vec Fint1{0.0, 0.0, 0.0};
            for (int n = 0; n < lookup.numBeads; n++) {
                const Bead& b2 = lookup.beadList[n];
                if (!(b1 == b2)) {
                    const Vec Fint2 = aziz.gradV(lookup.beadSep[n]);
                    for (int i = 0; i < 3; i++) Fint1[i] -= Fint2[i];
                    const Vec Fext2{0.0, 0.0, 0.0};
                    Vec Fint3{0.0, 0.0, 0.0};
                    for (int m = 0; m < lookup.numBeads; m++) {
                        const Bead& b3 = lookup.beadList[m];
                        if (!(b3 == b2) && !(b3 == b1)) {
                            Vec g = aziz.gradV(getSeparation(b2, b3));
                            for (int i = 0; i < 3; i++) Fint3[i] += g[i];
                        }
                    }
                    totF2 += dot3(Fint2, Fint2) + 2.0 * dot3(Fext2, Fint2) + 2.0 * dot3(Fint2, Fint3);
                }
            }
            totF2 += dot3(Fext1, Fext1) + 2.0 * dot3(Fext1, Fint1) + dot3(Fint1, Fint1);
        }
        return totF2;
    }
    // the correction of potentialAction(bead) (action.cpp:497-503): at shift 1 only
    double correction(const Bead& b) const {
        int eo = b[0] % 2;
        if ((shift == 1) && (gradVFactor[eo] > kEPS))
            return gradVFactor[eo] * tauShift() * tauShift() * tauShift() * lambda * gradVnnSquared(b);
        return 0.0;
    }
    // barePotentialAction(vector, values) (action.cpp:647-689): VFactor tau Vnn per
    // active bead, 0 for the others, in list order
    void bareActionList(const std::vector<Bead>& beadsIn, std::vector<double>& values) {
        values.assign(beadsIn.size(), 0.0);
        for (size_t i = 0; i < beadsIn.size(); ++i) {
            const Bead& b = beadsIn[i];
            if (!beadOn(b)) continue;
            double ext = factor(getState(b)) * 0.0;
            values[i] = VFactor[b[0] % 2] * tauShift() * Vnn(b, ext);
        }
    }
    // Virial pair terms (action.cpp:1446-1805; the free external potential has
    // gradV = grad2V = 0).  The T-matrix weights the transverse part with
    // dV = |gradV| (a magnitude, not the signed dV/dr) as the reference does.
    void tMatAdd(const Vec& rDiff, double rmag, double dV, double d2V, double tMat[3][3]) const {
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) {
                tMat[a][c] += rDiff[a] * rDiff[c] * d2V / (rmag * rmag) - rDiff[a] * rDiff[c] * dV / std::pow(rmag, 3);
                if (a == c) tMat[a][c] += dV / rmag;
            }
    }
    Vec sliceForce(const Bead& b1) const {
        Vec gVi{0.0, 0.0, 0.0};
        Bead b2{b1[0], 0};
        for (b2[1] = 0; b2[1] < numBeadsAtSlice[b1[0]]; b2[1]++)
            if (!(b1 == b2)) {
                Vec g = aziz.gradV(getSeparation(b1, b2));
                for (int i = 0; i < 3; i++) gVi[i] += g[i];
            }
        return gVi;   // + externalPtr->gradV(path(bead1)) = 0
    }
    // gVdotT = T gV for bead1 (the bead2 loop of rDOTgradUterm2, :1530-1566)
    Vec forceDotT(const Bead& b1) const {
        Vec gV{0.0, 0.0, 0.0};
        double tMat[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        Bead b2{b1[0], 0};
        for (b2[1] = 0; b2[1] < numBeadsAtSlice[b1[0]]; b2[1]++) {
            Vec rDiff = getSeparation(b1, b2);
            double rmag = std::sqrt(dot3(rDiff, rDiff));
            if (!(b1 == b2)) {
                Vec gVi = aziz.gradV(rDiff);
                double dVi = std::sqrt(dot3(gVi, gVi));
                double g2Vi = aziz.grad2V(rDiff);
                double dV = dVi + 0.0, d2V = g2Vi + 0.0;
                tMatAdd(rDiff, rmag, dV, d2V, tMat);
                for (int i = 0; i < 3; i++) gV[i] += gVi[i];
            }
        }
        Vec out{0.0, 0.0, 0.0};
        for (int r = 0; r < 3; r++) {
            double ip = 0.0;
            for (int c = 0; c < 3; c++) ip += tMat[r][c] * gV[c];
            out[r] += ip;
        }
        return out;
    }
    // pos1 - centre of the worldline window, put in the box (:1621-1645)
    Vec windowDelta(const Bead& b1, int W) const {
        Vec runMore{0, 0, 0}, runLess{0, 0, 0}, COM{0, 0, 0};
        const Vec pos1 = pos(b1);
        Bead nOld = b1, pOld = b1;
        for (int gamma = 0; gamma < W; gamma++) {
            Bead bn = next(b1, gamma), bp = prev(b1, gamma);
            Vec sm = getSeparation(bn, nOld), sl = getSeparation(bp, pOld);
            for (int i = 0; i < 3; i++) {
                runMore[i] += sm[i];
                runLess[i] += sl[i];
                COM[i] += (pos1[i] + runMore[i]) + (pos1[i] + runLess[i]);
            }
            nOld = bn;
            pOld = bp;
        }
        Vec delta;
        for (int i = 0; i < 3; i++) {
            COM[i] /= (2.0 * W);
            delta[i] = pos1[i] - COM[i];
        }
        box.putInBC(delta);
        return delta;
    }
    double rDOTgradUterm1(int slice) const {   // :1446-1482
        double rDotgV = 0.0;
        Bead b1{slice, 0};
        for (b1[1] = 0; b1[1] < numBeadsAtSlice[slice]; b1[1]++) rDotgV += dot3(sliceForce(b1), pos(b1));
        return VFactor[slice % 2] * tauShift() * rDotgV;
    }
    double rDOTgradUterm2(int slice) const {   // :1494-1574
        const int eo = slice % 2;
        double term2 = 0.0;
        if (gradVFactor[eo] > kEPS) {
            Bead b1{slice, 0};
            for (b1[1] = 0; b1[1] < numBeadsAtSlice[slice]; b1[1]++) term2 += dot3(forceDotT(b1), pos(b1));
            term2 *= (2.0 * gradVFactor[eo] * std::pow(tauShift(), 3) * lambda);
        }
        return term2;
    }
    double deltadotgradUterm1(int slice, int W) const {   // :1587-1653
        double rDotgV = 0.0;
        Bead b1{slice, 0};
        for (b1[1] = 0; b1[1] < numBeadsAtSlice[slice]; b1[1]++)
            rDotgV += dot3(sliceForce(b1), windowDelta(b1, W));
        return VFactor[slice % 2] * tauShift() * rDotgV;
    }
    double deltadotgradUterm2(int slice, int W) const {   // :1665-1762
        const int eo = slice % 2;
        double term2 = 0.0;
        if (gradVFactor[eo] > kEPS) {
            Bead b1{slice, 0};
            for (b1[1] = 0; b1[1] < numBeadsAtSlice[slice]; b1[1]++)
                term2 += dot3(forceDotT(b1), windowDelta(b1, W));
            term2 *= (2.0 * gradVFactor[eo] * std::pow(tauShift(), 3) * lambda);
        }
        return term2;
    }
    double virialKinCorrection(int slice) const {   // :1769-1786
        const int eo = slice % 2;
        double vKC = 0.0;
        if (gradVFactor[eo] > kEPS) {
            vKC += gradVSquared(slice);
            vKC *= gradVFactor[eo] * std::pow(tauShift(), 3) * lambda;
        }
        return vKC;
    }
    double secondDerivTau(int slice) const {   // :776-787
        const int eo = slice % 2;
        double d2U = 0.0;
        if (gradVFactor[eo] > kEPS) d2U += 6.0 * gradVFactor[eo] * tauShift() * lambda * gradVSquared(slice);
        return d2U;
    }
    // potentialActionCorrection(bead) (action.cpp:700-716): any shift
    double actionCorrection(const Bead& b) {
        int eo = b[0] % 2;
        if (gradVFactor[eo] > kEPS) {
            lookup.updateInteractionList(*this, b);
            return gradVFactor[eo] * tauShift() * tauShift() * tauShift() * lambda * gradVnnSquared(b);
        }
        return 0.0;
    }
    double actionBead(const Bead& b) {
        // action.cpp:487-514
        int eo = b[0] % 2;
        double bareU = VFactor[eo] * tauShift() * Vnn(b);
        double corU = beadOn(b) ? correction(b) : 0.0;
        return bareU + corU;
    }
    double actionList(const std::vector<Bead>& beadsIn) {
        // action.cpp:572-620
        std::vector<Bead> active;
        std::vector<double> ext;
        for (const auto& b : beadsIn)
            if (beadOn(b)) {
                active.push_back(b);
                ext.push_back(factor(getState(b)) * 0.0);
            }
        double totU = 0.0;
        for (size_t i = 0; i < active.size(); ++i) {
            int eo = active[i][0] % 2;
            double bareU = VFactor[eo] * tauShift() * Vnn(active[i], ext[i]);
            double corU = correction(active[i]);   // uses the list Vnn just built
            totU += bareU + corU;
        }
        return totU;
    }
    double actionRange(const Bead& start, const Bead& end) {
        // action.cpp:522-570: walk next() from start to end inclusive
        std::vector<Bead> list;
        Bead b = start;
        Bead stop = next(end);
        do {
            list.push_back(b);
            b = next(b);
        } while (!(b == stop));
        return actionList(list);
    }
    double totalAction() {
        // action.cpp:456-472
        double totU = 0.0;
        for (int slice = 0; slice < M; slice++) {
            int eo = slice % 2;
            auto v = Vslice(slice);
            double s = 0.0;
            s += v[0];
            s += v[1];
            totU += VFactor[eo] * tauShift() * s;
            if (gradVFactor[eo] > kEPS)
                totU += gradVFactor[eo] * tauShift() * tauShift() * tauShift() * lambda * gradVSquared(slice);
        }
        return totU;
    }
    void updateBead(const Bead& b, const Vec& p) {
        // src/path.cpp:197-200
        lookup.updateBead(b, p);
        beads[idx(b)] = p;
    }
    void initDefaultTopology() {
        // src/path.cpp:28-129 (PIMC: periodic links in imaginary time)
        on.assign(size_t(M) * nalloc, 1);
        numBeadsAtSlice.assign(M, nalloc);
        nextL.resize(size_t(M) * nalloc);
        prevL.resize(size_t(M) * nalloc);
        for (int i = 0; i < M; i++)
            for (int j = 0; j < nalloc; j++) {
                prevL[size_t(i) * nalloc + j] = Bead{i - 1, j};
                nextL[size_t(i) * nalloc + j] = Bead{i + 1, j};
            }
        for (int j = 0; j < nalloc; j++) {
            prevL[j] = Bead{M - 1, j};
            nextL[size_t(M - 1) * nalloc + j] = Bead{0, j};
        }
        head = tail = sp1 = sp2 = Bead{kXXX, kXXX};
    }
};

void Lookup::updateGrid(const State& st) {
    // (:314-349)
    std::fill(numLabels.begin(), numLabels.end(), 0);
    Bead b;
    for (b[0] = 0; b[0] < st.M; ++b[0])
        for (b[1] = 0; b[1] < st.nalloc; ++b[1])
            if (st.beadOn(b)) {
                grid[st.idx(b)] = gridIndex(st.pos(b));
                append(b);
            }
}

void Lookup::updateInteractionList(const State& st, const Bead& b1) {
    // (:464-516)
    numBeads = 0;
    IVec g1 = grid[st.idx(b1)];
    int c1 = cellNumber(g1);
    Bead b2{b1[0], 0};
    for (int n = 0; n < numNN; n++) {
        IVec g = gridNN[size_t(c1) * numNN + n];
        if (g[0] != kXXX && g[1] != kXXX && g[2] != kXXX) {
            int cell = cellNumber(g);
            int maxNL = NL(cell, b1[0]);
            for (int label = 0; label < maxNL; label++) {
                b2[1] = H(cell, b1[0], label);
                if (!(b1 == b2)) {
                    Vec sep = st.getSeparation(b2, b1);
                    if (dot3(sep, sep) < rc2) {
                        beadList[numBeads] = b2;
                        beadSep[numBeads] = sep;
                        numBeads++;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Diagonal moves (src/move.cpp): Bisection, Staging, CenterOfMass.
// ---------------------------------------------------------------------------
struct Moves {
    State* st = nullptr;
    int Mbar = 8, b = 3;
    bool varUpdates = true;
    int maxWind = 1;
    double comDelta = 0;
    double sqrtLambdaTau = 0, sqrt2LambdaTau = 0;
    std::vector<IVec> winding;
    std::vector<double> cumrho0;
    // scratch
    std::vector<Vec> originalPos, newPos;
    std::vector<char> include;
    Bead startBead{0, 0}, endBead{0, 0};

    void init(State* s, int Mbar_, int b_, bool var, int mw, double cd) {
        st = s;
        Mbar = Mbar_;
        b = b_;
        varUpdates = var;
        maxWind = mw;
        comDelta = cd;
        // MoveBase ctor (move.cpp:58-112)
        sqrtLambdaTau = std::sqrt(st->lambda * st->tau);
        sqrt2LambdaTau = std::sqrt(2.0) * sqrtLambdaTau;
        int nw1 = 2 * maxWind + 1;
        int numWind = static_cast<int>(std::floor(std::pow(1.0 * nw1, 3.0) + kEPS));
        cumrho0.assign(numWind, 0.0);
        winding.clear();
        for (int n = 0; n < numWind; n++) {
            IVec w{0, 0, 0};
            for (int i = 0; i < 3; i++) {
                int scale = 1;
                for (int j = i + 1; j < 3; j++) scale *= nw1;
                w[i] = (n / scale) % nw1;
                w[i] -= (w[i] > maxWind) * nw1;
            }
            winding.push_back(w);
        }
        // The reference's stable_sort comparator compares w1 with itself
        // (move.cpp:120-134) and is therefore always false: order is unchanged.
    }
    double rho0(const Vec& vel, int Mlen) const {
        // action.cpp rho0(dVec, int)
        double norm = std::pow(4.0 * M_PI * st->lambda * Mlen * st->tau, -0.5 * kNDIM);
        return (norm * std::exp(-dot3(vel, vel) / (4.0 * st->lambda * Mlen * st->tau)));
    }
    IVec sampleWindingSector(MT& rng, const Bead& s, const Bead& e, int stageLength) {
        // move.cpp:359-396
        const Vec& pe = st->pos(e);
        const Vec& ps = st->pos(s);
        Vec vel{pe[0] - ps[0], pe[1] - ps[1], pe[2] - ps[2]};
        cumrho0[0] = rho0(vel, stageLength);
        double totalrho0 = cumrho0[0];
        for (size_t n = 1; n < winding.size(); n++) {
            Vec vw;
            for (int i = 0; i < 3; i++) vw[i] = vel[i] + winding[n][i] * st->box.side[i];
            double c = rho0(vw, stageLength);
            totalrho0 += c;
            cumrho0[n] = cumrho0[n - 1] + c;
        }
        for (auto& c : cumrho0) c /= totalrho0;
        double u = rng.rand();
        size_t index = std::lower_bound(cumrho0.begin(), cumrho0.end(), u) - cumrho0.begin();
        return winding.at(index);
    }
    Vec newStagingPosition(MT& rng, const Bead& nb, const Bead& e, int stageLength, int k, IVec& wind) {
        // move.cpp:306-346 (winding form)
        double f1 = 1.0 * (stageLength - k - 1);
        double f2 = 1.0 / (1.0 * (stageLength - k));
        double sq = sqrt2LambdaTau * std::sqrt(f1 * f2);
        const Vec& np = st->pos(nb);
        const Vec& pe = st->pos(e);
        Vec r;
        for (int i = 0; i < 3; i++) r[i] = (pe[i] + st->box.side[i] * wind[i]) - np[i];
        for (int i = 0; i < 3; i++) r[i] *= f2;
        for (int i = 0; i < 3; i++) r[i] += np[i];
        for (int i = 0; i < 3; i++) r[i] = rng.randNorm(r[i], sq);
        for (int i = 0; i < 3; i++) {
            while (r[i] < -0.5 * st->box.side[i]) {
                wind[i]++;
                r[i] += st->box.side[i];
            }
            while (r[i] >= 0.5 * st->box.side[i]) {
                wind[i]--;
                r[i] -= st->box.side[i];
            }
        }
        st->box.putInBC(r);  // putInside == putInBC for a fully periodic prism
        return r;
    }
    Vec newBisectionPosition(MT& rng, const Bead& bi, int lshift) {
        // move.cpp:467-488
        double delta = sqrtLambdaTau * std::sqrt(1.0 * lshift);
        Bead nbi = st->prev(bi, lshift);
        Vec r = st->getSeparation(st->next(bi, lshift), nbi);
        for (int i = 0; i < 3; i++) r[i] *= 0.5;
        const Vec& np = st->pos(nbi);
        for (int i = 0; i < 3; i++) r[i] += np[i];
        for (int i = 0; i < 3; i++) r[i] = rng.randNorm(r[i], delta);
        st->box.putInBC(r);
        return r;
    }

    int bisection(MT& rng, double* info) {
        // move.cpp:1484-1661 (local action branch)
        if (info) info[0] = info[1] = info[2] = 0.0;
        if (st->numBeadsOn() == 0) return 0;
        int numLevels = b;
        if (varUpdates) numLevels = 1 + static_cast<int>(rng.randInt(b - 1));
        int numActive = static_cast<int>(std::floor(std::pow(2.0, 1.0 * numLevels) + kEPS)) - 1;
        if (st->trueNumParticles() == 0) return 0;
        startBead[0] = static_cast<int>(rng.randInt(st->M - 1));
        if (st->numBeadsAtSlice[startBead[0]] == 0) return 0;
        startBead[1] = static_cast<int>(rng.randInt(st->numBeadsAtSlice[startBead[0]] - 1));
        Bead bi = startBead;
        for (int k = 0; k < numActive + 1; k++) {
            if (!st->beadOn(bi) || (st->next(bi)[0] == kXXX && st->next(bi)[1] == kXXX)) return 0;
            bi = st->next(bi);
        }
        endBead = bi;
        if (info) {
            info[0] = 1.0;
            info[2] = numLevels;
        }
        include.assign(numActive, 1);
        originalPos.assign(numActive, Vec{0, 0, 0});
        newPos.assign(numActive, Vec{0, 0, 0});
        double oldDelta = 0.0;
        for (int level = numLevels; level > 0; level--) {
            int shift = static_cast<int>(std::floor(std::pow(2.0, 1.0 * (level - 1)) + kEPS));
            st->shift = shift;
            double oldAction = 0.0, newAction = 0.0;
            std::vector<Bead> inc, prevLevel;
            std::vector<int> incIdx, prevIdx;
            bi = st->next(startBead, shift);
            int k = 1;
            do {
                int n = k * shift - 1;
                if (include[n]) {
                    originalPos[n] = st->pos(bi);
                    inc.push_back(bi);
                    incIdx.push_back(n);
                } else if (level == 1) {
                    prevLevel.push_back(bi);
                    prevIdx.push_back(n);
                }
                ++k;
                bi = st->next(bi, shift);
            } while (!(bi == endBead));
            oldAction += st->actionList(inc);
            for (size_t i = 0; i < inc.size(); ++i) {
                int n = incIdx[i];
                newPos[n] = newBisectionPosition(rng, inc[i], shift);
                st->updateBead(inc[i], newPos[n]);
                include[n] = 0;
            }
            newAction += st->actionList(inc);
            if (level == 1) {
                newAction += st->actionList(prevLevel);
                for (size_t i = 0; i < prevLevel.size(); ++i)
                    st->updateBead(prevLevel[i], originalPos[prevIdx[i]]);
                oldAction += st->actionList(prevLevel);
                for (size_t i = 0; i < prevLevel.size(); ++i)
                    st->updateBead(prevLevel[i], newPos[prevIdx[i]]);
            }
            double delta = (newAction - oldAction);
            if (info) info[1] = delta;
            if (rng.rand() < std::exp(-delta + oldDelta)) {
                if (level == 1) {
                    st->shift = 1;
                    return 1;
                }
            } else {
                // undoMove (move.cpp:1686-1703)
                int kk = 0;
                Bead u = startBead;
                do {
                    u = st->next(u);
                    if (!include[kk]) st->updateBead(u, originalPos[kk]);
                    ++kk;
                } while (!(u == st->prev(endBead)));
                st->shift = 1;
                return 0;
            }
            oldDelta = delta;
        }
        st->shift = 1;
        return 0;
    }

    int staging(MT& rng, double* info) {
        // move.cpp:1315-1408
        if (info) info[0] = info[1] = info[2] = 0.0;
        if (st->numBeadsOn() == 0) return 0;
        if (st->trueNumParticles() == 0) return 0;
        int stageLength = Mbar;
        startBead[0] = static_cast<int>(rng.randInt(st->M - 1));
        if (st->numBeadsAtSlice[startBead[0]] == 0) return 0;
        startBead[1] = static_cast<int>(rng.randInt(st->numBeadsAtSlice[startBead[0]] - 1));
        if (varUpdates) stageLength = 2 + static_cast<int>(rng.randInt(Mbar - 2));
        Bead bi = startBead;
        for (int k = 0; k < stageLength; k++) {
            if (!st->beadOn(bi) || (st->next(bi)[0] == kXXX && st->next(bi)[1] == kXXX)) return 0;
            bi = st->next(bi);
        }
        endBead = bi;
        if (info) {
            info[0] = 1.0;
            info[2] = stageLength;
        }
        IVec wind = sampleWindingSector(rng, startBead, endBead, stageLength);
        double oldAction = st->actionRange(startBead, st->prev(endBead));
        originalPos.assign(stageLength, Vec{0, 0, 0});
        bi = startBead;
        int k = 0;
        do {
            bi = st->next(bi);
            originalPos[k] = st->pos(bi);
            st->updateBead(bi, newStagingPosition(rng, st->prev(bi), endBead, stageLength, k, wind));
            ++k;
        } while (!(bi == st->prev(endBead)));
        double newAction = st->actionRange(startBead, st->prev(endBead));
        if (info) info[1] = newAction - oldAction;
        if (rng.rand() < std::exp(-(newAction - oldAction))) return 1;
        k = 0;
        bi = startBead;
        do {
            bi = st->next(bi);
            st->updateBead(bi, originalPos[k]);
            ++k;
        } while (!(bi == st->prev(endBead)));
        return 0;
    }

    int com(MT& rng, double* info) {
        // move.cpp:1124-1259 (PIMC, no broken worldlines)
        if (info) info[0] = info[1] = info[2] = 0.0;
        if (st->numBeadsOn() == 0) return 0;
        int startSlice = 0;
        if (st->numBeadsAtSlice[startSlice] == 0) return 0;
        Bead first{startSlice, static_cast<int>(rng.randInt(st->numBeadsAtSlice[startSlice] - 1))};
        auto isNull = [](const Bead& b) { return b[0] == kXXX && b[1] == kXXX; };
        Bead s = first;
        if (!isNull(st->prev(s))) {
            do { s = st->prev(s); } while (!(st->prev(s) == first) && !isNull(st->prev(s)));
        }
        Bead e;
        if (st->prev(s) == first) {
            s = first;
            e = st->prev(s);
        } else {
            e = first;
            if (!isNull(st->next(e))) {
                do { e = st->next(e); } while (!isNull(st->next(e)));
            }
        }
        startBead = s;
        endBead = e;
        int wlLength = 0;
        Bead bi = s;
        do {
            ++wlLength;
            bi = st->next(bi);
        } while (!(bi == st->next(e)));
        if (wlLength > st->M) return 0;
        if (info) {
            info[0] = 1.0;
            info[2] = wlLength;
        }
        Vec shiftV;
        for (int i = 0; i < 3; i++) shiftV[i] = comDelta * (-0.5 + rng.rand());
        double oldAction = st->actionRange(s, e);
        bi = s;
        do {
            const Vec& p = st->pos(bi);
            Vec np{p[0] + shiftV[0], p[1] + shiftV[1], p[2] + shiftV[2]};
            st->box.putInBC(np);
            st->updateBead(bi, np);
            bi = st->next(bi);
        } while (!(bi == st->next(e)));
        double newAction = st->actionRange(s, e);
        if (info) info[1] = newAction - oldAction;
        if (rng.rand() < std::exp(-(newAction - oldAction))) return 1;
        bi = s;
        do {
            const Vec& p = st->pos(bi);
            Vec np{p[0] - shiftV[0], p[1] - shiftV[1], p[2] - shiftV[2]};
            st->box.putInBC(np);
            st->updateBead(bi, np);
            bi = st->next(bi);
        } while (!(bi == st->next(e)));
        return 0;
    }
};

std::vector<std::string> tokenize(const char* s) {
    std::istringstream iss(s ? s : "");
    std::vector<std::string> t;
    std::string w;
    while (iss >> w) t.push_back(w);
    return t;
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

void* orc_mt_create(uint32_t seed) { return new MT(seed); }
void orc_mt_destroy(void* mt) { delete static_cast<MT*>(mt); }
uint32_t orc_mt_randint(void* mt) { return static_cast<MT*>(mt)->randInt(); }
uint32_t orc_mt_randint_n(void* mt, uint32_t n) { return static_cast<MT*>(mt)->randInt(n); }
double orc_mt_rand(void* mt) { return static_cast<MT*>(mt)->rand(); }
double orc_mt_randexc(void* mt) { return static_cast<MT*>(mt)->randExc(); }
double orc_mt_randnorm(void* mt, double mean, double sd) { return static_cast<MT*>(mt)->randNorm(mean, sd); }
void orc_mt_save(void* mt, uint32_t* out) {
    MT* m = static_cast<MT*>(mt);
    for (int i = 0; i < MT::N; i++) out[i] = m->s[i];
    out[MT::N] = static_cast<uint32_t>(m->left);
}
void orc_mt_load(void* mt, const uint32_t* in) {
    MT* m = static_cast<MT*>(mt);
    for (int i = 0; i < MT::N; i++) m->s[i] = in[i];
    m->left = static_cast<int>(in[MT::N]);
    m->next = MT::N - m->left;
}

void* orc_aziz_create(int year, double max_sep, double rc, int build_derivs) {
    Aziz* a = new Aziz();
    a->init(year, max_sep, rc, build_derivs != 0);
    return a;
}
void orc_aziz_destroy(void* az) { delete static_cast<Aziz*>(az); }
double orc_aziz_valueV(void* az, double r) { return static_cast<Aziz*>(az)->valueV(r); }
double orc_aziz_valuedVdr(void* az, double r) { return static_cast<Aziz*>(az)->valuedVdr(r); }
double orc_aziz_valued2Vdr2(void* az, double r) { return static_cast<Aziz*>(az)->valued2Vdr2(r); }
double orc_aziz_V(void* az, const double* s) { return static_cast<Aziz*>(az)->V(Vec{s[0], s[1], s[2]}); }
void orc_aziz_gradV(void* az, const double* s, double* out) {
    Aziz* a = static_cast<Aziz*>(az);
    if (a->tdV.empty()) { out[0] = out[1] = out[2] = NAN; return; }
    Vec g = a->gradV(Vec{s[0], s[1], s[2]});
    out[0] = g[0]; out[1] = g[1]; out[2] = g[2];
}
double orc_aziz_grad2V(void* az, const double* s) {
    Aziz* a = static_cast<Aziz*>(az);
    if (a->td2V.empty()) return NAN;
    return a->grad2V(Vec{s[0], s[1], s[2]});
}
double orc_aziz_tailV(void* az) { return static_cast<Aziz*>(az)->tailV; }
double orc_aziz_dr(void* az) { return static_cast<Aziz*>(az)->dr; }
int orc_aziz_table_length(void* az) { return static_cast<Aziz*>(az)->tableLength; }
const double* orc_aziz_table(void* az, int which) {
    Aziz* a = static_cast<Aziz*>(az);
    const std::vector<double>& t = which == 0 ? a->tV : (which == 1 ? a->tdV : a->td2V);
    return t.empty() ? nullptr : t.data();
}
void orc_aziz_branch_bins(void* az, int32_t* out3) {
    Aziz* a = static_cast<Aziz*>(az);
    out3[0] = out3[1] = out3[2] = a->tableLength;
    double r = 0;
    bool f0 = false, f1 = false, f2 = false;
    for (int n = 0; n < a->tableLength; n++) {
        double x = r / a->rm;
        if (!f0 && !(x < kEPS)) { out3[0] = n; f0 = true; }
        if (!f1 && !(x < 0.01)) { out3[1] = n; f1 = true; }
        if (!f2 && !(x < a->D)) { out3[2] = n; f2 = true; }
        if (f0 && f1 && f2) break;
        r += a->dr;
    }
}
void orc_aziz_params(void* az, double* o) {
    Aziz* a = static_cast<Aziz*>(az);
    o[0] = a->epsilon; o[1] = a->rm; o[2] = a->A; o[3] = a->alpha; o[4] = a->beta;
    o[5] = a->D; o[6] = a->C6; o[7] = a->C8; o[8] = a->C10;
}

void* orc_state_create(double side, int M, int nalloc, double tau, double lambda, double rc, int year) {
    State* st = new State();
    st->box.init(side);
    st->M = M;
    st->nalloc = nalloc;
    st->tau = tau;
    st->lambda = lambda;
    st->rc = rc;
    st->aziz.init(year, st->box.maxSep, rc, false);
    st->lookup.setup(&st->box, M, nalloc, rc);
    st->beads.assign(size_t(M) * nalloc, Vec{0, 0, 0});
    st->initDefaultTopology();
    st->dSep = 0.5 * std::sqrt(3.0) * st->box.side[2] / (1.0 * kNPCFSEP);
    return st;
}
void orc_state_destroy(void* p) { delete static_cast<State*>(p); }
void* orc_state_aziz(void* p) { return &static_cast<State*>(p)->aziz; }
void orc_state_set_beads(void* p, const double* b) {
    State* st = static_cast<State*>(p);
    for (size_t i = 0; i < st->beads.size(); i++) st->beads[i] = Vec{b[3 * i], b[3 * i + 1], b[3 * i + 2]};
    st->initDefaultTopology();
    st->lookup.updateGrid(*st);
}
void orc_state_set_topology(void* p, const uint8_t* bead_on, const int32_t* num_beads, const int32_t* next,
                            const int32_t* prev, const int32_t* head, const int32_t* tail,
                            const int32_t* special1, const int32_t* special2) {
    State* st = static_cast<State*>(p);
    size_t n = size_t(st->M) * st->nalloc;
    for (size_t i = 0; i < n; i++) {
        st->on[i] = bead_on[i];
        st->nextL[i] = Bead{next[2 * i], next[2 * i + 1]};
        st->prevL[i] = Bead{prev[2 * i], prev[2 * i + 1]};
    }
    for (int m = 0; m < st->M; m++) st->numBeadsAtSlice[m] = num_beads[m];
    st->head = Bead{head[0], head[1]};
    st->tail = Bead{tail[0], tail[1]};
    st->sp1 = Bead{special1[0], special1[1]};
    st->sp2 = Bead{special2[0], special2[1]};
    st->lookup.updateGrid(*st);
}
void orc_state_get_beads(void* p, double* b) {
    State* st = static_cast<State*>(p);
    for (size_t i = 0; i < st->beads.size(); i++)
        for (int d = 0; d < 3; d++) b[3 * i + d] = st->beads[i][d];
}
void orc_state_set_action(void* p, int shift, double vf0, double vf1) {
    State* st = static_cast<State*>(p);
    st->shift = shift;
    st->VFactor[0] = vf0;
    st->VFactor[1] = vf1;
}
double orc_total_action(void* p) { return static_cast<State*>(p)->totalAction(); }
double orc_total_action_threads(void* p, int threads) {
    State* st = static_cast<State*>(p);
    if (threads < 1 || st->gradVFactor[0] > kEPS || st->gradVFactor[1] > kEPS) return std::nan("");
    return st->totalActionThreads(threads);
}
void orc_V_slice(void* p, int slice, double* out2, int32_t* hist) {
    State* st = static_cast<State*>(p);
    auto v = st->Vslice(slice);
    out2[0] = v[0];
    out2[1] = v[1];
    if (hist)
        for (int i = 0; i < kNPCFSEP; i++) hist[i] = st->sepHist[i];
}
double orc_Vnn_bead(void* p, int slice, int ptcl) { return static_cast<State*>(p)->Vnn(Bead{slice, ptcl}); }
double orc_Vnn_slice(void* p, int slice) { return static_cast<State*>(p)->VnnSlice(slice); }
double orc_action_bead(void* p, int slice, int ptcl) { return static_cast<State*>(p)->actionBead(Bead{slice, ptcl}); }
double orc_action_range(void* p, const int32_t* s, const int32_t* e) {
    return static_cast<State*>(p)->actionRange(Bead{s[0], s[1]}, Bead{e[0], e[1]});
}
double orc_action_list(void* p, const int32_t* b, int n) {
    std::vector<Bead> list;
    for (int i = 0; i < n; i++) list.push_back(Bead{b[2 * i], b[2 * i + 1]});
    return static_cast<State*>(p)->actionList(list);
}
void orc_update_bead(void* p, int slice, int ptcl, const double* pos3) {
    static_cast<State*>(p)->updateBead(Bead{slice, ptcl}, Vec{pos3[0], pos3[1], pos3[2]});
}
void orc_state_set_correction(void* p, double gvf0, double gvf1, int period) {
    State* st = static_cast<State*>(p);
    st->gradVFactor[0] = gvf0;
    st->gradVFactor[1] = gvf1;
    st->period = period < 1 ? 1 : period;
    if ((gvf0 > kEPS || gvf1 > kEPS) && st->aziz.tdV.empty())
        st->aziz.init(st->aziz.year, st->box.maxSep, st->rc, true);   // gradV needs lookupdVdr
}
double orc_gradV2_slice(void* p, int slice) { return static_cast<State*>(p)->gradVSquared(slice); }
void orc_bare_action_list(void* p, const int32_t* b, int n, double* out) {
    std::vector<Bead> list;
    for (int i = 0; i < n; i++) list.push_back(Bead{b[2 * i], b[2 * i + 1]});
    std::vector<double> v;
    static_cast<State*>(p)->bareActionList(list, v);
    for (int i = 0; i < n; i++) out[i] = v[i];
}
double orc_action_correction(void* p, int slice, int ptcl) {
    return static_cast<State*>(p)->actionCorrection(Bead{slice, ptcl});
}
void orc_left_pack(int M, int N, double* beads, int32_t* next, int32_t* prev, uint32_t* worm) {
    // Path::leftPack (path.cpp:145-189) on raw [M][N] arrays; links {slice, ptcl},
    // XXX = {-1,-1}
    auto at = [N](int m, int i) { return size_t(m) * N + i; };
    for (int m = 0; m < M; m++) {
        for (int i = 0; i < N; i++) {
            if (worm[at(m, i)]) continue;
            bool found = false;
            int j;
            for (j = i + 1; j < N; j++)
                if (worm[at(m, j)]) {
                    found = true;
                    break;
                }
            if (!found) continue;
            size_t b1 = at(m, i), b2 = at(m, j);
            for (int d = 0; d < 3; d++) beads[3 * b1 + d] = beads[3 * b2 + d];
            worm[b1] = 1;   // worm.addBead(bead1)
            worm[b2] = 0;   // worm.delBead(bead2)
            next[2 * b1] = next[2 * b2];
            next[2 * b1 + 1] = next[2 * b2 + 1];
            prev[2 * b1] = prev[2 * b2];
            prev[2 * b1 + 1] = prev[2 * b2 + 1];
            // next(prev(bead2)) = bead1; prev(next(bead2)) = bead1 (skipped where
            // the link is XXX: the reference would index with XXX there)
            int pm = prev[2 * b2], pp = prev[2 * b2 + 1];
            if (!(pm == -1 && pp == -1)) {
                next[2 * at(pm, pp)] = m;
                next[2 * at(pm, pp) + 1] = i;
            }
            int nm = next[2 * b2], np_ = next[2 * b2 + 1];
            if (!(nm == -1 && np_ == -1)) {
                prev[2 * at(nm, np_)] = m;
                prev[2 * at(nm, np_) + 1] = i;
            }
            next[2 * b2] = next[2 * b2 + 1] = -1;
            prev[2 * b2] = prev[2 * b2 + 1] = -1;
        }
    }
}
void orc_virial_terms(void* p, int W, double* out) {
    State* st = static_cast<State*>(p);
    if (st->aziz.tdV.empty()) st->aziz.init(st->aziz.year, st->box.maxSep, st->rc, true);
    for (int m = 0; m < st->M; m++) {
        double* o = out + 6 * (size_t)m;
        o[0] = st->rDOTgradUterm1(m);
        o[1] = st->rDOTgradUterm2(m);
        o[2] = st->deltadotgradUterm1(m, W);
        o[3] = st->deltadotgradUterm2(m, W);
        o[4] = st->virialKinCorrection(m);
        o[5] = st->secondDerivTau(m);
    }
}
void orc_virial_energy(void* p, int W, double mu, double* out) {
    // VirialEnergyEstimator::accumulate (estimator.cpp:1086-1251), one call
    State* st = static_cast<State*>(p);
    if (st->aziz.tdV.empty()) st->aziz.init(st->aziz.year, st->box.maxSep, st->rc, true);
    for (int i = 0; i < 19; i++) out[i] = 0.0;
    double totVop = 0.0, thermE = 0.0, totEcv = 0.0, Kcv = 0.0, T2 = 0.0, T3 = 0.0, T4 = 0.0, virKinTerm = 0.0;
    const int numParticles = st->trueNumParticles();
    const int nts = st->M;
    const double tau = st->tau, lambda = st->lambda;
    const double volume = st->box.volume;
    const double beta = 1.0 * nts * tau;
    const double tailV = (1.0 * numParticles * numParticles / volume) * st->aziz.tailV;
    const double thermTerm1 = (0.5 * kNDIM / tau) * numParticles;
    const double T1 = 0.5 * kNDIM * numParticles / (1.0 * W * tau);
    const double exchangeNorm = 1.0 / (4.0 * W * std::pow(tau, 2) * lambda * nts);
    double Pressure = kNDIM * numParticles;
    double P2 = 0.0, P3 = 0.0;
    Bead b1;
    for (int slice = 0; slice < nts; slice++) {
        b1[0] = slice;
        for (int ptcl = 0; ptcl < st->numBeadsAtSlice[slice]; ptcl++) {
            b1[1] = ptcl;
            Vec vel2 = st->getVelocity(b1);
            Bead nOld = b1;
            Vec vel1{0, 0, 0};
            for (int gamma = 1; gamma <= W; gamma++) {
                Bead bn = st->next(b1, gamma);
                Vec s = st->getSeparation(bn, nOld);
                for (int i = 0; i < 3; i++) vel1[i] += s[i];
                nOld = bn;
            }
            T2 -= dot3(vel1, vel2);
            thermE -= dot3(vel2, vel2);
        }
    }
    P2 = thermE;
    T2 *= exchangeNorm;
    P2 /= (2.0 * lambda * tau * nts);
    Pressure += P2;
    double T5 = 0.0;
    for (int slice = 0; slice < nts; slice++) {
        const int eo = slice % 2;
        T3 += st->deltadotgradUterm1(slice, W);
        T4 += st->deltadotgradUterm2(slice, W);
        // derivPotentialActionTau (action.cpp:751-765)
        auto v = st->Vslice(slice);
        double dU = st->VFactor[eo] * (v[0] + v[1]);
        if (st->gradVFactor[eo] > kEPS)
            dU += 3.0 * st->gradVFactor[eo] * st->tauShift() * st->tauShift() * lambda * st->gradVSquared(slice);
        T5 += dU;
        virKinTerm += st->virialKinCorrection(slice);
        if (eo == 0) totVop += v[0] + v[1];
        P3 += st->rDOTgradUterm1(slice) + st->rDOTgradUterm2(slice);
    }
    P3 *= (1.0 / (2.0 * nts));
    Pressure -= P3;
    Pressure /= (kNDIM * tau * volume);
    totVop /= (0.5 * nts);
    totVop += tailV;
    T3 /= (2.0 * beta);
    T4 /= (1.0 * beta);
    T5 /= (1.0 * nts);
    virKinTerm /= (0.5 * beta);
    thermE *= (1.0 / (4.0 * lambda * tau)) / (tau * nts);
    thermE += thermTerm1;
    thermE += T5;
    thermE += tailV;
    totEcv = T1 + T2 + T3 + T4 + T5 + tailV;
    Kcv = T1 + T2 + T3 + T4 + virKinTerm;
    double dEdB = (-1.0 * T1 - 2.0 * T2 + 2.0 * T4) / tau;
    for (int slice = 0; slice < nts; slice++) dEdB += st->secondDerivTau(slice) / (1.0 * nts);
    dEdB *= beta * beta / (1.0 * nts);
    // columns: K_op K_cv V_op V_cv E E_mu K_op/N K_cv/N V_op/N V_cv/N E/N
    //          EEcv*Beta^2 Ecv*Beta dEdB CvCov1 CvCov2 CvCov3 E_th P
    out[0] += totEcv - totVop;
    out[1] += Kcv;
    out[2] += totVop;
    out[3] += totEcv - Kcv;
    out[4] += totEcv;
    out[5] += totEcv - mu * numParticles;
    if (numParticles > 0) {
        out[6] += (totEcv - totVop) / (1.0 * numParticles);
        out[7] += Kcv / (1.0 * numParticles);
        out[8] += totVop / (1.0 * numParticles);
        out[9] += (totEcv - Kcv) / (1.0 * numParticles);
        out[10] += totEcv / (1.0 * numParticles);
    }
    out[11] += totEcv * thermE * beta * beta;
    out[12] += totEcv * beta;
    out[13] += dEdB;
    out[14] += totEcv * thermE * beta * beta * totEcv * beta;
    // estIndex["cVCov2"]: a key the constructor never registered; std::map
    // inserts it with index 0, so this lands on K_op and CvCov2 stays 0
    out[0] += totEcv * beta * dEdB;
    out[16] += totEcv * thermE * beta * beta * dEdB;
    out[17] += thermE;
    out[18] += Pressure;
}
double orc_gradVnn2(void* p, int slice, int ptcl) {
    State* st = static_cast<State*>(p);
    Bead b{slice, ptcl};
    st->lookup.updateInteractionList(*st, b);
    return st->gradVnnSquared(b);
}

void orc_energy(void* p, double mu, double* out) {
    // EnergyEstimator::accumulate (estimator.cpp:940-1029), PIMC, primitive/local action
    State* st = static_cast<State*>(p);
    int N = st->trueNumParticles();
    int nts = st->M;  // endSlice - startSlice
    double tailV = (1.0 * N * N / st->box.volume) * st->aziz.tailV;
    double fourLambdaTauInv = 1.0 / (4.0 * st->lambda * st->tau);
    double kinNorm = fourLambdaTauInv / (st->tau * nts);
    double classicalKinetic = (0.5 * kNDIM / st->tau) * N;
    double totK = 0.0;
    Bead b;
    for (int slice = 0; slice < nts; slice++) {
        b[0] = slice;
        int nb = st->numBeadsAtSlice[slice];
        for (int ptcl = 0; ptcl < nb; ptcl++) {
            b[1] = ptcl;
            Vec v = st->getVelocity(b);
            totK -= dot3(v, v);
        }
    }
    totK *= kinNorm;
    double t1 = 0.0, t2 = 0.0;
    std::array<double, 2> totVop{0.0, 0.0};
    const int period = st->period;
    const double tau = st->tauShift();
    for (int slice = 0; slice < nts; slice++) {
        int eo = slice % 2;
        const bool cor = st->gradVFactor[eo] > kEPS;
        const double g2 = cor ? st->gradVSquared(slice) : 0.0;
        // derivPotentialActionLambda (action.cpp:795-802)
        t1 += 1.0 * (cor ? st->gradVFactor[eo] * tau * tau * tau * g2 : 0.0);
        auto v = st->Vslice(slice);
        double s = 0.0;
        s += v[0];
        s += v[1];
        // derivPotentialActionTau (action.cpp:751-767)
        double dU = st->VFactor[eo] * s;
        if (cor) dU += 3.0 * st->gradVFactor[eo] * tau * tau * st->lambda * g2;
        t2 += 1.0 * dU;
        if (!(slice % period)) {
            auto w = st->Vslice(slice);
            totVop[0] += 1.0 * w[0];
            totVop[1] += 1.0 * w[1];
        }
    }
    t1 *= st->lambda / (st->tau * nts);
    t2 /= 1.0 * nts;
    totVop[0] /= (nts / period);
    totVop[1] /= (nts / period);
    totK += (classicalKinetic + t1);
    double totV = t2 - t1 + tailV;
    totVop[1] += tailV;
    out[0] = totK;
    out[1] = totV;
    out[2] = totVop[0];
    out[3] = totVop[1];
    out[4] = totK + totV;
    out[5] = totK + totV - mu * N;   // E_mu (estimator.cpp:1011)
    out[6] = N > 0 ? totK / (1.0 * N) : 0.0;
    out[7] = N > 0 ? totV / (1.0 * N) : 0.0;
    out[8] = N > 0 ? (totK + totV) / (1.0 * N) : 0.0;
}

void orc_ssf(void* p, const double* q3, int nq, double* out) {
    // estimator.cpp:3705-3737
    State* st = static_cast<State*>(p);
    int N = st->trueNumParticles();
    for (int iq = 0; iq < nq; iq++) {
        Vec q{q3[3 * iq], q3[3 * iq + 1], q3[3 * iq + 2]};
        double sf = 0.0;
        Bead b1, b2;
        for (b1[0] = 0; b1[0] < st->M; b1[0]++) {
            b2[0] = b1[0];
            for (b1[1] = 0; b1[1] < st->numBeadsAtSlice[b1[0]]; b1[1]++) {
                sf += 1.0;
                for (b2[1] = b1[1] + 1; b2[1] < st->numBeadsAtSlice[b2[0]]; b2[1]++)
                    sf += 2 * std::cos(dot3(st->getSeparation(b1, b2), q));
            }
        }
        out[iq] = sf / N;
    }
}

void orc_isf(void* p, const double* q3, int nq, double* out) {
    // estimator.cpp:3923-3961
    State* st = static_cast<State*>(p);
    int N = st->trueNumParticles();
    int M = st->M;
    std::vector<double> isf(size_t(nq) * M, 0.0);
    for (int i = 0; i < nq; i++) {
        Vec q{q3[3 * i], q3[3 * i + 1], q3[3 * i + 2]};
        Bead b1, b2;
        for (b1[0] = 0; b1[0] < M; b1[0]++)
            for (int tausep = 0; tausep < M; tausep++) {
                b2[0] = (b1[0] + tausep) % M;
                for (b1[1] = 0; b1[1] < st->numBeadsAtSlice[b1[0]]; b1[1]++) {
                    double lq1 = dot3(q, st->pos(b1));
                    for (b2[1] = 0; b2[1] < st->numBeadsAtSlice[b2[0]]; b2[1]++) {
                        double lq2 = dot3(q, st->pos(b2));
                        isf[size_t(i) * M + tausep] +=
                            (std::cos(lq1) * std::cos(lq2) + std::sin(lq1) * std::sin(lq2));
                    }
                }
            }
    }
    for (size_t k = 0; k < isf.size(); k++) out[k] = isf[k] / N;
}

void orc_grid_dims(void* p, int32_t* n3, int32_t* cap) {
    State* st = static_cast<State*>(p);
    for (int i = 0; i < 3; i++) n3[i] = st->lookup.nGrid[i];
    *cap = st->lookup.cap;
}
void orc_grid_dump(void* p, int32_t* grid, int32_t* beadLabel, int32_t* numLabels, int32_t* hash) {
    State* st = static_cast<State*>(p);
    Lookup& L = st->lookup;
    for (size_t i = 0; i < L.grid.size(); i++)
        for (int d = 0; d < 3; d++) grid[3 * i + d] = L.grid[i][d];
    for (size_t i = 0; i < L.beadLabel.size(); i++) beadLabel[i] = L.beadLabel[i];
    for (size_t i = 0; i < L.numLabels.size(); i++) numLabels[i] = L.numLabels[i];
    for (size_t i = 0; i < L.hash.size(); i++) hash[i] = L.hash[i];
}
void orc_grid_nn(void* p, int32_t* out) {
    State* st = static_cast<State*>(p);
    Lookup& L = st->lookup;
    for (size_t i = 0; i < L.gridNN.size(); i++)
        for (int d = 0; d < 3; d++) out[3 * i + d] = L.gridNN[i][d];
}
int orc_interaction_list(void* p, int slice, int ptcl, int32_t* list2, double* sep3) {
    State* st = static_cast<State*>(p);
    st->lookup.updateInteractionList(*st, Bead{slice, ptcl});
    for (int n = 0; n < st->lookup.numBeads; n++) {
        if (list2) {
            list2[2 * n] = st->lookup.beadList[n][0];
            list2[2 * n + 1] = st->lookup.beadList[n][1];
        }
        if (sep3)
            for (int d = 0; d < 3; d++) sep3[3 * n + d] = st->lookup.beadSep[n][d];
    }
    return st->lookup.numBeads;
}

int orc_qvectors(const double* side3, const char* type, const char* input, double* out3, int max_out) {
    // EstimatorBase::getQVectors (estimator.cpp:439-570), text inputs only
    std::string t(type ? type : "");
    std::vector<std::string> tok = tokenize(input);
    std::vector<Vec> qs;
    if (tok.empty()) return -1;
    if (t == "int" || t == "float") {
        for (size_t i = 0; i + 3 <= tok.size(); i += 3) {
            Vec q;
            for (int j = 0; j < 3; ++j) {
                if (t == "int") q[j] = (2.0 * M_PI / side3[j]) * std::stoi(tok[i + j]);
                else q[j] = std::stof(tok[i + j]);
            }
            qs.push_back(q);
        }
    } else if (t == "max_int" || t == "max_float") {
        IVec qmi{0, 0, 0}, qi{0, 0, 0};
        Vec qmax{0, 0, 0}, q{0, 0, 0};
        if (tok.size() < 3) return -1;
        for (int i = 0; i < 3; ++i) {
            if (t == "max_int") {
                qmi[i] = std::abs(std::stoi(tok[i]));
                qmax[i] = qmi[i] * 2.0 * M_PI / side3[i];
            } else {
                qmax[i] = std::stof(tok[i]);
            }
        }
        double qmag_max = std::sqrt(dot3(qmax, qmax));
        if (t == "max_float")
            for (int i = 0; i < 3; ++i) qmi[i] = 1 + static_cast<int>(qmag_max * side3[i] / 2.0 / M_PI);
        int n_q = 1;
        for (int i = 0; i < 3; ++i) {
            n_q *= 2 * qmi[i] + 1;
            qi[i] = -qmi[i];
            q[i] = qi[i] * 2.0 * M_PI / side3[i];
        }
        if (std::sqrt(dot3(q, q)) <= qmag_max) qs.push_back(q);
        int pos = 2, count = 0;
        while (count < n_q - 1) {
            if (qi[pos] == qmi[pos]) {
                qi[pos] = -qmi[pos];
                pos -= 1;
            } else {
                qi[pos] += 1;
                for (int i = 0; i < 3; ++i) q[i] = qi[i] * 2.0 * M_PI / side3[i];
                if (std::sqrt(dot3(q, q)) <= qmag_max) qs.push_back(q);
                count += 1;
                pos = 2;
            }
        }
    } else {
        return -1;
    }
    int n = static_cast<int>(qs.size());
    for (int i = 0; i < n && i < max_out; i++)
        for (int d = 0; d < 3; d++) out3[3 * i + d] = qs[i][d];
    return n;
}

void orc_gen_worldlines(uint32_t seed, double side, int M, int N, double lambda, double tau, int kind,
                        double* beads) {
    MT rng(seed);
    Box box;
    box.init(side);
    std::vector<Vec> b(size_t(M) * N);
    // slice 0: Prism::randPosition (container.cpp:155-160)
    for (int p = 0; p < N; p++)
        for (int i = 0; i < 3; i++) b[p][i] = side * (-0.5 + rng.randExc());
    if (kind == 0) {
        double sd = std::sqrt(2.0 * lambda * tau);
        for (int m = 1; m < M; m++)
            for (int p = 0; p < N; p++) {
                Vec r;
                for (int i = 0; i < 3; i++) r[i] = rng.randNorm(b[size_t(m - 1) * N + p][i], sd);
                box.putInBC(r);
                b[size_t(m) * N + p] = r;
            }
    } else {
        // closed staging bridge from slice 0 back to slice 0 through M-1 beads
        double s2lt = std::sqrt(2.0) * std::sqrt(lambda * tau);
        int stageLength = M;
        for (int p = 0; p < N; p++) {
            const Vec end = b[p];
            for (int k = 0; k < M - 1; k++) {
                const Vec& nb = b[size_t(k) * N + p];
                double f1 = 1.0 * (stageLength - k - 1);
                double f2 = 1.0 / (1.0 * (stageLength - k));
                double sq = s2lt * std::sqrt(f1 * f2);
                Vec r{end[0] - nb[0], end[1] - nb[1], end[2] - nb[2]};
                box.putInBC(r);
                for (int i = 0; i < 3; i++) r[i] *= f2;
                for (int i = 0; i < 3; i++) r[i] += nb[i];
                for (int i = 0; i < 3; i++) r[i] = rng.randNorm(r[i], sq);
                box.putInBC(r);
                b[size_t(k + 1) * N + p] = r;
            }
        }
    }
    for (size_t i = 0; i < b.size(); i++)
        for (int d = 0; d < 3; d++) beads[3 * i + d] = b[i][d];
}

void* orc_moves_create(void* st, int Mbar, int b, int var_updates, int max_wind, double com_delta) {
    Moves* m = new Moves();
    m->init(static_cast<State*>(st), Mbar, b, var_updates != 0, max_wind, com_delta);
    return m;
}
void orc_moves_destroy(void* mv) { delete static_cast<Moves*>(mv); }
int orc_move_bisection(void* mv, void* mt, double* info) {
    return static_cast<Moves*>(mv)->bisection(*static_cast<MT*>(mt), info);
}
int orc_move_staging(void* mv, void* mt, double* info) {
    return static_cast<Moves*>(mv)->staging(*static_cast<MT*>(mt), info);
}
int orc_move_com(void* mv, void* mt, double* info) {
    return static_cast<Moves*>(mv)->com(*static_cast<MT*>(mt), info);
}

}  // extern "C"
