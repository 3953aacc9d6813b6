// api_suite.cpp -- GPU checks of the drop-in header include/tensor.hpp.
//
// Inputs and expected values come from data: the reference suite's
// constants (GPUtils test/testTensor.cu, cited X:LINE) are read at run time
// from tests/golden/reference_vectors.txt, which tests/golden/make_golden.py
// writes from tests/golden_values.py; everything else is computed on the
// host from those inputs (transposes, products, slices, sums).  Each check
// runs for the element types the reference instantiates.  The suite runs
// with two Session streams so that stream routing is exercised too.
//
//   api_suite [substring]     runs the scenarios whose name contains it
// Exit status: number of failed scenarios.
#include <cmath>
#include <cstdio>
#include <functional>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "tensor.hpp"

/* ------------------------------------------------------------ harness --- */

namespace suite {

struct Scenario {
    std::string name;
    std::function<void()> body;
};

std::vector<Scenario> &registry() {
    static std::vector<Scenario> all;
    return all;
}

int g_failures = 0;  // failed checks in the running scenario

void report(bool ok, const std::string &what, const char *file, int line) {
    if (ok) return;
    ++g_failures;
    std::fprintf(stderr, "    check failed at %s:%d: %s\n", file, line, what.c_str());
}

void close(double got, double want, double tol, const char *expr, const char *file, int line) {
    const bool ok = std::fabs(got - want) <= tol;
    if (!ok) {
        char msg[256];
        std::snprintf(msg, sizeof msg, "%s = %.17g, expected %.17g (tol %.1e)", expr, got, want, tol);
        report(false, msg, file, line);
    }
}

template <typename Exc, typename F>
void throws(F &&f, const char *expr, const char *file, int line) {
    bool caught = false;
    try {
        f();
    } catch (const Exc &) {
        caught = true;
    }
    report(caught, std::string("no expected exception from ") + expr, file, line);
}

/** Calls f(T{}) once for every listed element type. */
template <typename... Ts, typename F>
void forTypes(F &&f) {
    (f(Ts{}), ...);
}

/** Named vectors of tests/golden/reference_vectors.txt. */
class Vectors {
public:
    Vectors() {
        const char *dir = std::getenv("GPUTILS_GOLDEN_DIR");
        const std::string path = std::string(dir ? dir : "tests/golden") + "/reference_vectors.txt";
        std::ifstream in(path);
        if (!in) throw std::runtime_error("cannot open " + path);
        std::string name;
        size_t count;
        while (in >> name >> count) {
            std::vector<double> v(count);
            for (double &x : v) in >> x;
            m_data[name] = std::move(v);
        }
    }
    template <typename T = double>
    std::vector<T> get(const std::string &name) const {
        auto it = m_data.find(name);
        if (it == m_data.end()) throw std::runtime_error("no vector " + name);
        return std::vector<T>(it->second.begin(), it->second.end());
    }
    double scalar(const std::string &name) const { return get(name).at(0); }

private:
    std::map<std::string, std::vector<double>> m_data;
};

const Vectors &data() {
    static Vectors v;
    return v;
}

}  // namespace suite

#define SCENARIO(id)                                                                        \
    static void id();                                                                       \
    static const bool id##_added = (suite::registry().push_back({#id, id}), true);          \
    static void id()
#define CHECK(cond) suite::report((cond), #cond, __FILE__, __LINE__)
#define CHECK_CLOSE(got, want, tol) suite::close((double)(got), (double)(want), (double)(tol), #got, __FILE__, __LINE__)
#define CHECK_THROWS(Exc, ...) suite::throws<Exc>([&] { (void)(__VA_ARGS__); }, #__VA_ARGS__, __FILE__, __LINE__)

using suite::data;
using suite::forTypes;

/* ------------------------------------------------------- host helpers --- */

namespace host {

template <typename T>
constexpr double tol() {
    return std::is_same_v<T, float> ? 1e-4 : 1e-10;  // the reference's PRECISION_LOW / _HIGH (X:5-6)
}

template <typename T>
std::vector<T> download(const DTensor<T> &t) {
    std::vector<T> v;
    t.download(v);
    return v;
}

/** Element (i, j, l) of column-major data of shape (m, n, *). */
template <typename T>
T at(const std::vector<T> &v, size_t m, size_t n, size_t i, size_t j, size_t l = 0) {
    return v[i + m * (j + n * l)];
}

/** C_l = A_l B_l for column-major batches (A m x p, B p x n). */
template <typename T>
std::vector<double> matmul(const std::vector<T> &A, const std::vector<T> &B, size_t m, size_t p, size_t n, size_t batch) {
    std::vector<double> C(m * n * batch, 0.0);
    for (size_t l = 0; l < batch; l++)
        for (size_t j = 0; j < n; j++)
            for (size_t q = 0; q < p; q++)
                for (size_t i = 0; i < m; i++)
                    C[i + m * (j + n * l)] += (double)A[i + m * (q + p * l)] * (double)B[q + p * (j + n * l)];
    return C;
}

template <typename T>
std::vector<T> transpose(const std::vector<T> &A, size_t m, size_t n, size_t batch) {
    std::vector<T> out(A.size());
    for (size_t l = 0; l < batch; l++)
        for (size_t j = 0; j < n; j++)
            for (size_t i = 0; i < m; i++) out[j + n * (i + m * l)] = A[i + m * (j + n * l)];
    return out;
}

/** Row-major (per matrix) data of shape (m, n, batch) as column-major. */
template <typename T>
std::vector<T> fromRowMajor(const std::vector<T> &rm, size_t m, size_t n, size_t batch) {
    return transpose(rm, n, m, batch);  // a row-major m x n matrix is a column-major n x m one
}

template <typename T>
double maxAbsDiff(const std::vector<T> &a, const std::vector<double> &b) {
    double d = 0;
    for (size_t i = 0; i < a.size(); i++) d = std::max(d, std::fabs((double)a[i] - b[i]));
    return d;
}

/** Seeded uniform values in [lo, hi] (splitmix64 bits). */
inline std::vector<double> uniform(size_t n, double lo, double hi, uint64_t seed) {
    std::vector<double> v(n);
    for (auto &x : v) {
        uint64_t z = (seed += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        x = lo + (hi - lo) * ((double)(z >> 11) * 0x1.0p-53);
    }
    return v;
}

/** Batch of SPD matrices M M^T + n I from seeded M. */
inline std::vector<double> spd(size_t n, size_t batch, uint64_t seed) {
    const auto M = uniform(n * n * batch, -1, 1, seed);
    auto A = matmul(M, transpose(M, n, n, batch), n, n, n, batch);
    for (size_t l = 0; l < batch; l++)
        for (size_t i = 0; i < n; i++) A[i + n * (i + n * l)] += (double)n;
    return A;
}

}  // namespace host

/* ============================================== storage and views ===== */

SCENARIO(zeroed_construction) {  // X:30-46
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        DTensor<T> z(2, 3, 4, true);
        CHECK(z.numRows() == 2 && z.numCols() == 3 && z.numMats() == 4);
        const auto v = host::download(z);
        CHECK(std::all_of(v.begin(), v.end(), [](T x) { return x == T(0); }));
    });
}

SCENARIO(row_and_column_major_input) {  // X:52-106
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        const size_t m = 3, n = 2, k = 2;
        std::vector<T> iota(m * n * k);
        std::iota(iota.begin(), iota.end(), T(1));
        const auto asColumnMajor = host::fromRowMajor(iota, m, n, k);
        DTensor<T> cm(asColumnMajor, m, n, k, columnMajor), rm(iota, m, n, k, rowMajor);
        CHECK(host::download(cm) == asColumnMajor);
        CHECK(host::download(rm) == asColumnMajor);
        // per-matrix uploads into axis-2 views of a larger tensor
        DTensor<T> whole(m, n, k);
        for (size_t l = 0; l < k; l++) {
            DTensor<T> view(whole, 2, l, l);
            std::vector<T> part(iota.begin() + l * m * n, iota.begin() + (l + 1) * m * n);
            view.upload(part, rowMajor);
        }
        CHECK(host::download(whole) == asColumnMajor);
    });
}

SCENARIO(random_tensor) {  // X:112-127
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        auto r = DTensor<T>::createRandomTensor(20, 40, 60, -1, 1);
        CHECK(r.numRows() == 20 && r.numCols() == 40 && r.numMats() == 60);
        const auto v = host::download(r);
        CHECK(std::all_of(v.begin(), v.end(), [](T x) { return x >= T(-1) && x <= T(1); }));
    });
}

SCENARIO(file_round_trips) {  // X:133-186: text and .bt, saved then parsed
    for (const char *name : {"round_trip.dtensor", "round_trip.bt"}) {
        forTypes<float, double>([&](auto tag) {
            using T = decltype(tag);
            for (int rep = 0; rep < 3; rep++) {
                auto r = DTensor<T>::createRandomTensor(20, 40, 6, -1, 1);
                r.saveToFile(name);
                auto back = DTensor<T>::parseFromFile(name);
                CHECK(back.numRows() == 20 && back.numCols() == 40 && back.numMats() == 6);
                CHECK(host::download(back) == host::download(r));  // both forms round-trip exactly
            }
        });
    }
    CHECK_THROWS(std::invalid_argument, DTensor<double>::parseFromFile("no_such_file.bt", rowMajor));  // X:209-214
    CHECK_THROWS(std::invalid_argument, DTensor<double>::parseFromFile("no_such_file.txt", rowMajor));
}

SCENARIO(bt_file_from_python_writer) {  // X:193-202: B(i, j, l) = 1 + l + 2j + 6i
    const char *dir = std::getenv("GPUTILS_GOLDEN_DIR");
    auto b = DTensor<double>::parseFromFile(std::string(dir ? dir : "tests/golden") + "/b_d.bt");
    CHECK(b.numRows() == 3 && b.numCols() == 3 && b.numMats() == 2);
    const auto v = host::download(b);
    for (size_t i = 0; i < 3; i++)
        for (size_t j = 0; j < 3; j++)
            for (size_t l = 0; l < 2; l++) CHECK(host::at(v, 3, 3, i, j, l) == 1.0 + l + 2.0 * j + 6.0 * i);
}

SCENARIO(construction_copy_and_move) {  // X:221-280
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("T234A");
        DTensor<T> t(a, 2, 3, 4);
        CHECK(t.numEl() == 24 && t.numMats() == 4);
        DTensor<T> copy(t);
        CHECK(copy.raw() != t.raw() && host::download(copy) == a);
        T *storage = copy.raw();
        DTensor<T> moved(std::move(copy));
        CHECK(moved.raw() == storage && moved.numEl() == 24);
        CHECK(copy.numEl() == 0 && copy.raw() == nullptr);
    });
    forTypes<int *, double *>([](auto tag) {  // pointer element types allocate and move too
        using T = decltype(tag);
        DTensor<T> z(2, 3, 4, true);
        DTensor<T> moved(std::move(z));
        DTensor<T> fromTemporary(DTensor<T>{100, 10, 1000});
        CHECK(z.numEl() == 0 && moved.numEl() == 24 && fromTemporary.numEl() == 1000000);
    });
}

SCENARIO(slices_view_the_parent) {  // X:287-355, X:860-893
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("T234A");
        DTensor<T> t(a, 2, 3, 4);
        DTensor<T> mats(t, 2, 1, 3);  // matrices 1..3
        CHECK(mats.numRows() == 2 && mats.numCols() == 3 && mats.numMats() == 3);
        CHECK(mats.raw() == t.raw() + 6);
        CHECK(mats.ptrMatrices() == t.ptrMatrices() + 1);
        CHECK(host::download(mats) == std::vector<T>(a.begin() + 6, a.end()));
        DTensor<T> cols(t, 1, 1, 2);  // columns 1..2 of matrix 0
        CHECK(cols.numRows() == 2 && cols.numCols() == 2 && cols.numMats() == 1 && cols.ptrMatrices() == nullptr);
        CHECK(host::download(cols) == std::vector<T>(a.begin() + 2, a.begin() + 6));
        DTensor<T> elems(t, 0, 2, 3);  // flat elements 2..3
        CHECK(elems.numRows() == 2 && elems.numCols() == 1 && elems.ptrMatrices() == nullptr);
        CHECK(host::download(elems) == std::vector<T>(a.begin() + 2, a.begin() + 4));
        CHECK_THROWS(std::invalid_argument, DTensor<T>(t, 2, 3, 4));  // past the last matrix
        CHECK_THROWS(std::invalid_argument, DTensor<T>(t, 1, 2, 3));
        CHECK_THROWS(std::invalid_argument, DTensor<T>(t, 0, 20, 24));
        CHECK_THROWS(std::invalid_argument, DTensor<T>(t, 2, 2, 1));
    });
}

SCENARIO(element_access_and_device_copies) {  // X:361-404, X:653-666
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("T234A");
        DTensor<T> t(2, 3, 4);
        t.upload(a);
        for (size_t l = 0; l < 4; l++)
            for (size_t j = 0; j < 3; j++)
                for (size_t i = 0; i < 2; i++) CHECK(t(i, j, l) == host::at(a, 2, 3, i, j, l));
        CHECK_THROWS(std::out_of_range, t(0, 0, 4));
        DTensor<T> target(2, 3, 5, true);
        DTensor<T> lastFour(target, 2, 1, 4);
        t.deviceCopyTo(lastFour);
        std::vector<T> want(6, T(0));
        want.insert(want.end(), a.begin(), a.end());
        CHECK(host::download(target) == want);
        DTensor<T> small(2, 3, 1);
        CHECK_THROWS(std::invalid_argument, t.deviceCopyTo(small));
    });
}

SCENARIO(reshape_keeps_elements) {  // X:411-448
    const size_t before = Session::getInstance().totalAllocatedBytes();
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        const size_t m = 5, n = 10, k = 3;
        auto a = DTensor<T>::createRandomTensor(m, n, k, -1, 1);
        const auto original = host::download(a);
        const size_t shapes[][3] = {{m, k, n}, {k, n, m}, {k * n, m, 1}, {m, k * n, 1}, {m * k * n, 1, 1}};
        for (const auto &s : shapes) {
            a.reshape(s[0], s[1], s[2]);
            CHECK(a.numRows() == s[0] && a.numCols() == s[1] && a.numMats() == s[2]);
            CHECK(host::download(a) == original);
        }
        CHECK(a(m * n * k - 1) == original.back() && a(0) == original.front());
        CHECK_THROWS(std::invalid_argument, a.reshape(7, 7, 7));
    });
    CHECK(Session::getInstance().totalAllocatedBytes() == before);  // exact accounting
}

SCENARIO(reshaped_slices_update_the_parent) {  // X:454-482
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("T234A"), b = data().get<T>("T234B");
        DTensor<T> A(a, 2, 3, 4), B(b, 2, 3, 4);
        DTensor<T> aTail(A, 2, 1, 3), bTail(B, 2, 1, 3);
        aTail.reshape(2, 9, 1);
        bTail.reshape(2, 9, 1);
        aTail += bTail;
        std::vector<double> want(a.begin(), a.end());
        for (size_t e = 6; e < 24; e++) want[e] += b[e];
        CHECK(host::maxAbsDiff(host::download(A), want) == 0);
        // a slice reshaped to more matrices gets a pointer table of its new shape
        DTensor<T> view(A, 2, 0, 1);
        view.reshape(2, 1, 6);
        CHECK(view.ptrMatrices() != nullptr && view.ptrMatrices() != A.ptrMatrices());
        std::vector<T *> table(6);
        gpuErrChk(gputils_memcpy(table.data(), view.ptrMatrices(), 6 * sizeof(T *), 1));
        for (size_t l = 0; l < 6; l++) CHECK(table[l] == A.raw() + 2 * l);
    });
}

SCENARIO(assignment_aliases) {  // X:672-688
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        DTensor<T> t(data().get<T>("T234A"), 2, 3, 4);
        DTensor<T> alias;
        alias = t;
        CHECK(alias.raw() == t.raw() && alias.ptrMatrices() == t.ptrMatrices());
        CHECK(alias.numRows() == 2 && alias.numCols() == 3 && alias.numMats() == 4);
        // an owner that becomes an alias frees what it held
        const size_t before = Session::getInstance().totalAllocatedBytes();
        DTensor<T> owner(100, 100, 2);
        owner = t;
        CHECK(Session::getInstance().totalAllocatedBytes() == before);
        // a view into the target's own storage keeps that storage alive
        DTensor<T> big(data().get<T>("T234A"), 2, 3, 4);
        big = DTensor<T>(big, 2, 1, 1);
        CHECK((host::download(big) == std::vector<T>{7, 8, 9, 8, 7, 10}));
    });
}

SCENARIO(bytes_are_counted) {  // X:955-992
    forTypes<float, double, int>([](auto tag) {
        using T = decltype(tag);
        const size_t before = Session::getInstance().totalAllocatedBytes();
        const size_t m = 10, n = 10, k = 20;
        auto *A = new DTensor<T>(m, n, k);
        CHECK(Session::getInstance().totalAllocatedBytes() == before + m * n * k * sizeof(T) + k * sizeof(T *));
        delete A;
        CHECK(Session::getInstance().totalAllocatedBytes() == before);
    });
}

/* ================================================= tensor algebra ===== */

SCENARIO(reductions) {  // X:488-573
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("T234A"), b = data().get<T>("T234B"), d = data().get<T>("T234AMB");
        double dot = 0, sq = 0, l1 = 0, dmax = 0, dmin = 1e300;
        for (size_t e = 0; e < a.size(); e++) {
            dot += (double)a[e] * b[e];
            sq += (double)a[e] * a[e];
            l1 += std::fabs((double)a[e]);
            dmax = std::max(dmax, std::fabs((double)d[e]));
            dmin = std::min(dmin, std::fabs((double)d[e]));
        }
        CHECK(dot == data().scalar("DOT_AB") && l1 == data().scalar("SUMABS_A"));  // the data agree with the goldens
        DTensor<T> va(a, a.size()), vb(b, b.size()), A(a, 2, 3, 4), B(b, 2, 3, 4), D(d, 2, 3, 4);
        CHECK_CLOSE(va.dotF(vb), dot, 0);
        CHECK_CLOSE(A.dotF(B), dot, 0);
        CHECK_CLOSE(A.normF(), data().scalar("NORMF_A"), host::tol<T>());
        CHECK_CLOSE(A.normF(), std::sqrt(sq), host::tol<T>());
        CHECK_CLOSE(A.sumAbs(), l1, 0);
        CHECK_CLOSE(D.maxAbs(), data().scalar("MAXABS_AMB"), 0);
        CHECK_CLOSE(D.minAbs(), dmin, 0);
        CHECK_CLOSE(D.maxAbs(), dmax, 0);
        CHECK_THROWS(std::invalid_argument, A.dotF(vb));
    });
}

SCENARIO(elementwise_arithmetic) {  // X:694-822
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("T234A"), b = data().get<T>("T234B");
        std::vector<double> sum(24), diff(24), triple(24);
        for (size_t e = 0; e < 24; e++) {
            sum[e] = (double)a[e] + b[e];
            diff[e] = (double)a[e] - b[e];
            triple[e] = 3.0 * a[e];
        }
        CHECK(host::maxAbsDiff(data().get<T>("T234APB"), sum) == 0);
        CHECK(host::maxAbsDiff(data().get<T>("T234AMB"), diff) == 0);
        DTensor<T> A(a, 2, 3, 4), B(b, 2, 3, 4);
        DTensor<T> scaled(A);
        scaled *= T(3);
        CHECK(host::maxAbsDiff(host::download(scaled), triple) == 0);
        CHECK(host::maxAbsDiff(host::download(T(3) * A), triple) == 0);
        CHECK(host::maxAbsDiff(host::download(A + B), sum) == 0);
        CHECK(host::maxAbsDiff(host::download(A - B), diff) == 0);
        DTensor<T> acc(A);
        acc += B;
        CHECK(host::maxAbsDiff(host::download(acc), sum) == 0);
        acc -= B;
        acc -= B;
        CHECK(host::maxAbsDiff(host::download(acc), diff) == 0);
        CHECK(host::download(A) == a);  // the operands are untouched
    });
}

SCENARIO(batched_products) {  // X:828-853
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("ADDAB_A"), b = data().get<T>("ADDAB_B");
        const auto want = host::matmul(a, b, 2, 3, 2, 3);
        CHECK(host::maxAbsDiff(data().get<T>("ADDAB_C"), want) == 0);
        DTensor<T> A(a, 2, 3, 3), B(b, 3, 2, 3), C(2, 2, 3, true);
        C.addAB(A, B);
        CHECK(host::maxAbsDiff(host::download(C), want) == 0);
        DTensor<T> P = A * B;
        CHECK(host::maxAbsDiff(host::download(P), want) == 0);
        // beta C + alpha A B
        C.addAB(A, B, T(2), T(-1));
        CHECK(host::maxAbsDiff(host::download(C), want) == 0);
    });
    // split-BF16 products (no reference counterpart): within 2^-15 k |A||B|
    const size_t m = 64, p = 64, n = 64, k = 50;
    const auto av = host::uniform(m * p * k, -1, 1, 7), bv = host::uniform(p * n * k, -1, 1, 8);
    std::vector<float> af(av.begin(), av.end()), bf(bv.begin(), bv.end());
    const auto want = host::matmul(af, bf, m, p, n, k);
    DTensor<float> A(af, m, p, k), B(bf, p, n, k), C(m, n, k), Cx(m, n, k);
    C.addAB(A, B);
    Cx.addAB(A, B, 1.0f, 0.0f, GemmMode::bf16x3);
    CHECK(host::maxAbsDiff(host::download(C), want) <= 1e-5 * p);
    CHECK(host::maxAbsDiff(host::download(Cx), want) <= std::ldexp(1.0, -15) * p);
    DTensor<double> D(2, 2, 1);
    CHECK_THROWS(std::invalid_argument, D.addAB(D, D, 1.0, 0.0, GemmMode::bf16x3));
}

SCENARIO(row_extraction) {  // X:899-926
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("GETROWS_A");
        DTensor<T> A(a, 3, 3, 2);
        for (size_t l = 0; l < 2; l++)
            for (size_t r0 = 0; r0 < 3; r0++)
                for (size_t r1 = r0; r1 < 3; r1++) {
                    DTensor<T> rows = A.getRows(r0, r1, l);
                    CHECK(rows.numRows() == r1 - r0 + 1 && rows.numCols() == 3);
                    std::vector<T> want;
                    for (size_t j = 0; j < 3; j++)
                        for (size_t i = r0; i <= r1; i++) want.push_back(host::at(a, 3, 3, i, j, l));
                    CHECK(host::download(rows) == want);
                }
    });
}

SCENARIO(transposes) {  // X:933-950
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("TR_A");
        CHECK(host::transpose(a, 3, 2, 2) == data().get<T>("TR_AT"));
        DTensor<T> A(a, 3, 2, 2);
        DTensor<T> At = A.tr();
        CHECK(At.numRows() == 2 && At.numCols() == 3 && At.numMats() == 2);
        CHECK(host::download(At) == host::transpose(a, 3, 2, 2));
    });
}

/* ======================================================= Cholesky ===== */

SCENARIO(cholesky_single_matrix) {  // X:1218-1276
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("CHOL_A"), rhs = data().get<T>("CHOL_B");
        DTensor<T> A(a, 3, 3, 1), L(A);
        CholeskyFactoriser<T> chol(L);
        chol.factorise();
        CHECK(chol.info()(0) == 0);
        const auto lowRm = data().get<T>("CHOL_L_ROWMAJOR");
        const auto low = host::fromRowMajor(lowRm, 3, 3, 1);
        for (size_t j = 0; j < 3; j++)
            for (size_t i = j; i < 3; i++) CHECK_CLOSE(L(i, j), host::at(low, 3, 3, i, j), host::tol<T>());
        DTensor<T> b(rhs, 3), x(b);
        chol.solve(x);
        const auto xs = data().get<T>("CHOL_X");
        for (size_t i = 0; i < 3; i++) CHECK_CLOSE(x(i), xs[i], host::tol<T>());
        DTensor<T> r = A * x;
        r -= b;
        CHECK(r.normF() < host::tol<T>());
    });
}

SCENARIO(cholesky_batches) {  // X:1282-1413: factor + solve, and solve from a given factor
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("CHOL_A"), rhs = data().get<T>("CHOL_B"), xs = data().get<T>("CHOL_X");
        const auto low = host::fromRowMajor(data().get<T>("CHOL_L_ROWMAJOR"), 3, 3, 1);
        std::vector<T> a2(a), low2(low), rhs2(rhs);
        a2.insert(a2.end(), a.begin(), a.end());
        low2.insert(low2.end(), low.begin(), low.end());
        rhs2.insert(rhs2.end(), rhs.begin(), rhs.end());
        for (bool given : {false, true}) {
            DTensor<T> A(a2, 3, 3, 2), F(given ? low2 : a2, 3, 3, 2);
            CholeskyBatchFactoriser<T> chol(F, given);
            chol.factorise();  // a no-op for a given factor
            for (size_t l = 0; l < 2; l++)
                for (size_t j = 0; j < 3; j++)
                    for (size_t i = j; i < 3; i++) CHECK_CLOSE(F(i, j, l), host::at(low, 3, 3, i, j), host::tol<T>());
            DTensor<T> b(rhs2, 3, 1, 2), x(b);
            chol.solve(x);
            for (size_t l = 0; l < 2; l++)
                for (size_t i = 0; i < 3; i++) CHECK_CLOSE(x(i, 0, l), xs[i], host::tol<T>());
            DTensor<T> r = A * x;
            r -= b;
            CHECK(r.normF() < host::tol<T>());
            CHECK(chol.info()(0, 0, 0) == 0 && chol.info()(0, 0, 1) == 0);
        }
    });
}

SCENARIO(cholesky_preconditions_and_failures) {  // T:2117, T:2163-2167; LAPACK info convention
    DTensor<double> notSquare(3, 2, 2);
    CHECK_THROWS(std::invalid_argument, CholeskyBatchFactoriser<double>(notSquare));
    DTensor<double> A(3, 3, 2, true);
    CholeskyBatchFactoriser<double> unfactorised(A);
    DTensor<double> b(3, 1, 2, true), tall(4, 1, 2, true);
    CHECK_THROWS(std::logic_error, unfactorised.solve(b));
    CholeskyBatchFactoriser<double> given(A, true);
    CHECK_THROWS(std::invalid_argument, given.solve(tall));
    for (size_t n : {2, 5, 40}) {  // indefinite leading minor of order 2
        std::vector<double> a(n * n, 0.0);
        for (size_t i = 0; i < n; i++) a[i + n * i] = 1.0;
        a[1] = a[n] = 2.0;
        DTensor<double> M(a, n, n, 1);
        CholeskyFactoriser<double> chol(M);
        chol.factorise();
        CHECK(chol.info()(0) == 2);
    }
}

SCENARIO(mixed_precision_cholesky) {  // no reference counterpart; beside T:2098-2197
    for (size_t n : {24, 100, 160}) {
        const size_t k = 37, nrhs = 3;
        const auto a = host::spd(n, k, 100 + n);
        const auto bv = host::uniform(n * nrhs * k, -1, 1, 200 + n);
        DTensor<double> A(a, n, n, k), B(bv, n, nrhs, k), X(n, nrhs, k);
        MixedCholeskyBatchSolver solver(A, 1e-12);
        solver.solve(B, X);
        DTensor<double> F(A), Xref(B);
        CholeskyBatchFactoriser<double> plain(F);
        plain.factorise();
        plain.solve(Xref);
        const auto x = host::download(X), xr = host::download(Xref);
        double scale = 0, diff = 0;
        for (size_t e = 0; e < x.size(); e++) {
            scale = std::max(scale, std::fabs(xr[e]));
            diff = std::max(diff, std::fabs(x[e] - xr[e]));
        }
        CHECK(diff <= 1e-10 * scale);
        const auto it = host::download(solver.iterations()), info = host::download(solver.info());
        for (size_t l = 0; l < k; l++) {
            CHECK(info[l] == 0);
            CHECK(n > 128 ? it[l] == -1 : it[l] >= 0);
        }
        DTensor<double> wrong(n + 1, nrhs, k);
        CHECK_THROWS(std::invalid_argument, solver.solve(B, wrong));
        CHECK_THROWS(std::invalid_argument, solver.solve(B, B));
    }
}

/* ========================================== least squares and QR ===== */

SCENARIO(batched_least_squares) {  // X:1012-1039
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const auto a = data().get<T>("LS1_A"), b = data().get<T>("LS1_B");
        DTensor<T> A(a, 2, 2, 3), work(A), B(b, 2, 1, 3), x(B);
        work.leastSquaresBatched(x);
        DTensor<T> r(2, 1, 3);
        r.addAB(A, x);
        r -= B;
        CHECK(r.normF() < host::tol<T>());
    });
    DTensor<double> A(3, 2, 2), fat(2, 3, 2), b4(4, 1, 2), b2(3, 2, 2), b1(3, 1, 1), bf(2, 1, 2);
    CHECK_THROWS(std::invalid_argument, A.leastSquaresBatched(b4));  // T:1344-1351
    CHECK_THROWS(std::invalid_argument, A.leastSquaresBatched(b2));
    CHECK_THROWS(std::invalid_argument, A.leastSquaresBatched(b1));
    CHECK_THROWS(std::invalid_argument, fat.leastSquaresBatched(bf));
}

SCENARIO(ragged_least_squares) {  // no reference counterpart: uneven sizes beside T:1340-1394
    const std::vector<int32_t> rows = {16, 300, 64, 512, 40, 9}, cols = {8, 64, 64, 50, 1, 9};
    RaggedLeastSquares plan(rows, cols);
    CHECK(plan.numSystems() == rows.size());
    const auto av = host::uniform(plan.elementsA(), -1, 1, 11), bv = host::uniform(plan.elementsB(), -1, 1, 12);
    DTensor<double> A(av, av.size()), B(bv, bv.size());
    plan.tune(A, B, 2);
    CHECK(host::download(B) == bv);  // tuning leaves b alone
    plan.solve(A, B, 2);
    const auto x = host::download(B);
    for (size_t s = 0; s < rows.size(); s++) {  // each system against the uniform batched solver
        const size_t m = rows[s], n = cols[s];
        std::vector<double> as(av.begin() + plan.offsetA(s), av.begin() + plan.offsetA(s) + m * n);
        std::vector<double> bs(bv.begin() + plan.offsetB(s), bv.begin() + plan.offsetB(s) + m);
        DTensor<double> As(as, m, n, 1), bsol(bs, m, 1, 1);
        As.leastSquaresBatched(bsol);
        const auto want = host::download(bsol);
        double scale = 0, diff = 0;
        for (size_t i = 0; i < n; i++) {
            scale = std::max(scale, std::fabs(want[i]));
            diff = std::max(diff, std::fabs(x[plan.offsetB(s) + i] - want[i]));
        }
        CHECK(diff <= 1e-9 * std::max(scale, 1.0));
    }
    CHECK_THROWS(std::invalid_argument, RaggedLeastSquares({10, 20}, {11, 5}));  // n > m
    DTensor<double> shortA(10);
    CHECK_THROWS(std::invalid_argument, plan.solve(shortA, B));
}

SCENARIO(qr_factorisation) {  // X:1433-1490
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        for (size_t m : {4, 20}) {
            const size_t n = 3;
            const auto av = host::uniform(m * n, -100, 100, 30 + m);
            DTensor<T> A(std::vector<T>(av.begin(), av.end()), m, n), work(A);
            QRFactoriser<T> qr(work);
            qr.factorise();
            CHECK(qr.info()(0) == 0);
            DTensor<T> Q(m, n), R(n, n, 1, true), QR(m, n);
            qr.getQR(Q, R);
            QR.addAB(Q, R);
            QR -= A;
            CHECK(QR.normF() < host::tol<T>() * 100);  // |A| ~ 100
            DTensor<T> Qt = Q.tr();
            DTensor<T> QtQ = Qt * Q;
            const auto g = host::download(QtQ);
            for (size_t j = 0; j < n; j++)
                for (size_t i = 0; i < n; i++) CHECK_CLOSE(g[i + n * j], i == j ? 1.0 : 0.0, host::tol<T>());
        }
    });
}

SCENARIO(qr_least_squares) {  // X:1496-1535
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const size_t m = 4, n = 3;
        DTensor<T> A(data().get<T>("QRLS_A_ROWMAJOR"), m, n, 1, rowMajor), work(A);
        DTensor<T> b(data().get<T>("QRLS_B"), m), xFull(b);
        DTensor<T> x(xFull, 0, 0, n - 1);
        QRFactoriser<T> qr(work);
        qr.factorise();
        qr.leastSquares(xFull);
        CHECK(qr.info()(0) == 0);
        DTensor<T> r(m);
        r.addAB(A, x);
        r -= b;
        CHECK_CLOSE(r.normF(), data().scalar("QRLS_RESIDUAL"), host::tol<T>());
    });
    DTensor<double> fat(2, 3), batch(3, 2, 2), A(4, 3), Qbad(4, 2), R(3, 3);
    CHECK_THROWS(std::invalid_argument, QRFactoriser<double>(fat));  // T:1811-1813
    CHECK_THROWS(std::invalid_argument, QRFactoriser<double>(batch));
    QRFactoriser<double> qr(A);
    CHECK_THROWS(std::invalid_argument, qr.getQR(Qbad, R));  // T:1933-1936
}

/* ============================================================ SVD ===== */

SCENARIO(svd_singular_values) {  // X:1058-1108
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        DTensor<T> B(data().get<T>("SVD8_B"), 8, 3);
        const auto before = host::download(B);
        Svd<T> svd(B, true, false);
        svd.factorise();
        CHECK(svd.info()(0) == 0);
        const auto s = data().get("SVD8_S01");
        CHECK_CLOSE(svd.singularValues()(0), s[0], host::tol<T>());
        CHECK_CLOSE(svd.singularValues()(1), s[1], host::tol<T>());
        CHECK(host::download(B) == before);
        // the accessors hand out the same storage every time
        CHECK(&svd.rightSingularVectors() == &svd.rightSingularVectors());
        CHECK(&svd.singularValues() == &svd.singularValues());
        auto u1 = svd.leftSingularVectors(), u2 = svd.leftSingularVectors();
        CHECK(u1.has_value() && u2.has_value() && *u1 == *u2 && (*u1)->raw() == (*u2)->raw());
        Svd<T> noU(B);
        CHECK(!noU.leftSingularVectors().has_value());
    });
    DTensor<double> fat(2, 3);
    CHECK_THROWS(std::invalid_argument, Svd<double>(fat));  // T:1518-1522
}

SCENARIO(svd_vectors_of_several_matrices) {  // X:1117-1171
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const double eps = std::is_same_v<T, float> ? 1e-3 : 1e-10;
        DTensor<T> A(data().get<T>("SVD3_A"), 3, 2, 3);
        Svd<T> svd(A, true);
        svd.factorise();
        const auto v = host::download(svd.rightSingularVectors()), s = host::download(svd.singularValues());
        const auto u = host::download(**svd.leftSingularVectors());
        const auto vw = data().get("SVD3_V4"), sw = data().get("SVD3_S"), uw = data().get("SVD3_U");
        for (size_t i = 0; i < vw.size(); i++) CHECK_CLOSE(v[i], vw[i], eps);
        for (size_t i = 0; i < sw.size(); i++) CHECK_CLOSE(s[i], sw[i], eps);
        for (size_t i = 0; i < uw.size(); i++) CHECK_CLOSE(u[i], uw[i], eps);
    });
}

SCENARIO(svd_rank) {  // X:1175-1199
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        DTensor<T> A(data().get<T>("RANK_A"), 4, 3, 3);
        Svd<T> svd(A);
        svd.factorise();
        const auto want = data().get("RANK_EXPECTED");
        for (int rep = 0; rep < 2; rep++) {  // recomputed, not accumulated (T:1602-1607 adds)
            const auto r = host::download(svd.rank(T(host::tol<T>())));
            for (size_t l = 0; l < 3; l++) CHECK(r[l] == (unsigned)want[l]);
        }
    });
}

SCENARIO(svd_wide_matrices) {  // tall matrices with more than 64 columns
    for (size_t n : {65, 96, 128}) {
        const size_t m = n + 7, k = 3;
        DTensor<double> A(host::uniform(m * n * k, -1, 1, n), m, n, k);
        Svd<double> svd(A, true);
        svd.factorise();
        const auto s = host::download(svd.singularValues());
        const auto vt = host::download(svd.rightSingularVectors()), u = host::download(**svd.leftSingularVectors());
        const auto a = host::download(A);
        for (size_t l = 0; l < k; l++) {  // A = U(:, 0:n) diag(s) V^T
            CHECK(svd.info()(0, 0, l) == 0);
            double err = 0;
            for (size_t j = 0; j < n; j++)
                for (size_t i = 0; i < m; i++) {
                    double acc = 0;
                    for (size_t q = 0; q < n; q++)
                        acc += u[i + m * (q + m * l)] * s[q + n * l] * vt[q + n * (j + n * l)];
                    err = std::max(err, std::fabs(acc - a[i + m * (j + n * l)]));
                }
            CHECK(err < 1e-10);
            for (size_t q = 1; q < n; q++) CHECK(s[q + n * l] <= s[q - 1 + n * l]);
        }
    }
}

/* ====================================================== Nullspace ===== */

SCENARIO(nullspace_bases) {  // X:1555-1592
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        DTensor<T> A(data().get<T>("NS_A"), 3, 4, 5);
        Nullspace<T> ns(A);
        DTensor<T> N = ns.nullspace();
        CHECK(N.numMats() == 5);
        for (size_t l = 0; l < 5; l++) {
            DTensor<T> Nl(N, 2, l, l), Al(A, 2, l, l);
            DTensor<T> AN = Al * Nl;
            CHECK(AN.normF() < host::tol<T>());
            DTensor<T> Nt = Nl.tr();
            DTensor<T> NtN = Nt * Nl;
            const auto g = host::download(NtN);
            CHECK_CLOSE(g[0], 1.0, host::tol<T>());
            for (size_t j = 0; j < 4; j++)
                for (size_t i = 0; i < 4; i++)
                    if (i != j) CHECK_CLOSE(g[i + 4 * j], 0.0, host::tol<T>());
        }
    });
}

SCENARIO(nullspace_of_full_rank_is_zero) {  // X:1598-1618
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        DTensor<T> A(data().get<T>("NS_TRIVIAL_ROWMAJOR"), 3, 3, 2, rowMajor);
        Nullspace<T> ns(A);
        CHECK(ns.nullspace().normF() == 0);
    });
}

SCENARIO(nullspace_projection) {  // X:1624-1663
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const size_t m = 3, n = 7;
        DTensor<T> A(m, n, 1);
        A.upload(data().get<T>("PROJ_A_ROWMAJOR"), rowMajor);
        Nullspace<T> ns(A);
        DTensor<T> N = ns.nullspace();
        DTensor<T> x(data().get<T>("PROJ_X"), n), px(x);
        ns.project(px);
        DTensor<T> Apx(m, 1, 1, true);
        Apx.addAB(A, px);
        CHECK(Apx.normF() < host::tol<T>());
        DTensor<T> other(data().get<T>("PROJ_OTHER"), n);
        DTensor<T> y = N * other;
        DTensor<T> d1 = y - px, d2 = px - x;
        CHECK(d1.dotF(d2) < host::tol<T>());
    });
}

SCENARIO(projection_replayed_from_a_graph) {  // repeated project through StreamGraph
    const size_t m = 3, n = 7;
    DTensor<double> A(m, n, 1);
    A.upload(data().get("PROJ_A_ROWMAJOR"), rowMajor);
    Nullspace<double> ns(A);
    const auto x = data().get("PROJ_X");
    DTensor<double> eager(x, n), b(x, n);
    ns.project(eager);
    StreamGraph g(b.streamIdx());
    g.capture([&] { ns.project(b); });
    for (int rep = 0; rep < 3; ++rep) {
        b.upload(x);
        g.launch();
        Session::getInstance().synchronizeAllStreams();
        CHECK(host::download(b) == host::download(eager));
    }
}

SCENARIO(nullspace_of_wide_matrices) {  // fat matrices with more than 64 rows
    const size_t m = 80, n = 120, k = 2, r = 70;
    // rank-r A = X Y from seeded factors
    const auto X = host::uniform(m * r * k, -1, 1, 41), Y = host::uniform(r * n * k, -1, 1, 42);
    const auto a = host::matmul(X, Y, m, r, n, k);
    DTensor<double> A(a, m, n, k);
    Nullspace<double> ns(A);
    DTensor<double> N = ns.nullspace();
    for (size_t l = 0; l < k; l++) {
        DTensor<double> Nl(N, 2, l, l), Al(A, 2, l, l);
        DTensor<double> AN = Al * Nl;
        CHECK(AN.normF() < 1e-9);
        DTensor<double> Nt = Nl.tr();
        DTensor<double> NtN = Nt * Nl;
        const auto g = host::download(NtN);
        for (size_t j = 0; j < n; j++)
            for (size_t i = 0; i < n; i++) CHECK_CLOSE(g[i + n * j], (i == j && j < n - r) ? 1.0 : 0.0, 1e-10);
    }
}

/* ========================================================= Givens ===== */

SCENARIO(givens_rotations) {  // X:580-641
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const size_t m = 10, n = 6;
        std::vector<T> v(m * n);
        std::iota(v.begin(), v.end(), T(1));
        const T c = T(0.1), s = std::sqrt(T(1) - c * c);
        {  // columns 1 and 4
            DTensor<T> a(v, m, n, 1);
            const T minus_s = s;
            a.applyRightGivensRotation(1, 4, &c, &minus_s);
            const auto got = host::download(a);
            for (size_t i = 0; i < m; i++)
                for (size_t j = 0; j < n; j++) {
                    double want = host::at(v, m, n, i, j);
                    if (j == 1) want = c * (double)host::at(v, m, n, i, 1) + minus_s * (double)host::at(v, m, n, i, 4);
                    if (j == 4) want = c * (double)host::at(v, m, n, i, 4) - minus_s * (double)host::at(v, m, n, i, 1);
                    CHECK_CLOSE(host::at(got, m, n, i, j), want, host::tol<T>() * 100);
                }
        }
        {  // rows 1 and 9
            DTensor<T> a(v, m, n, 1);
            const T minus_s = -s;
            a.applyLeftGivensRotation(1, 9, &c, &minus_s);
            const auto got = host::download(a);
            for (size_t j = 0; j < n; j++)
                for (size_t i = 0; i < m; i++) {
                    double want = host::at(v, m, n, i, j);
                    if (i == 1) want = c * (double)host::at(v, m, n, 1, j) + minus_s * (double)host::at(v, m, n, 9, j);
                    if (i == 9) want = c * (double)host::at(v, m, n, 9, j) - minus_s * (double)host::at(v, m, n, 1, j);
                    CHECK_CLOSE(host::at(got, m, n, i, j), want, host::tol<T>() * 100);
                }
        }
        DTensor<T> batch(3, 3, 2);
        CHECK_THROWS(std::invalid_argument, batch.applyRightGivensRotation(0, 1, &c, &s));
    });
}

SCENARIO(givens_annihilation) {  // X:1684-1729
    forTypes<float, double>([](auto tag) {
        using T = decltype(tag);
        const size_t m = 10, n = 6;
        std::vector<T> v(m * n);
        std::iota(v.begin(), v.end(), T(1));
        DTensor<T> a(v, m, n, 1);
        GivensAnnihilator<T> ga(a);
        for (size_t k = 1; k < m; k++)
            for (size_t j = 0; j < n; j++) {
                ga.annihilate(0, k, j);
                CHECK_CLOSE(a(k, j), 0.0, host::tol<T>());
            }
        DTensor<T> b(v, m, n, 1);
        ga.setMatrix(b);
        ga.annihilate(0, 1, 2);
        const double eps = std::is_same_v<T, float> ? 1e-4 : 1e-13;
        CHECK_CLOSE(b(1, 2), 0.0, eps);
        CHECK_CLOSE(b(0, 0), data().scalar("GIVENS_A02_ETC_A00"), eps);
        CHECK_CLOSE(b(0, 3), data().scalar("GIVENS_A02_ETC_A03"), eps * 10);
        CHECK_CLOSE(b(1, 3), data().scalar("GIVENS_A02_ETC_A13"), eps);
        CHECK_THROWS(std::invalid_argument, ga.annihilate(0, m, 0));
        CHECK_THROWS(std::invalid_argument, ga.annihilate(0, 1, n));  // j is checked too (T:2277 drops it)
    });
}

/* ======================================================== streams ===== */

SCENARIO(work_on_a_second_stream) {  // Session::setStreams, setStreamIdx (T:139-141, T:663-670)
    CHECK(Session::getInstance().numStreams() == 2);
    const size_t n = 48, k = 64;
    const auto a = host::spd(n, k, 5), bv = host::uniform(n * 2 * k, -1, 1, 6);
    const auto g1 = host::uniform(64 * 64 * 512, -1, 1, 7);
    // reference results, everything on stream 0
    DTensor<double> F0(a, n, n, k), X0(bv, n, 2, k), P0(n, 2, k);
    CholeskyBatchFactoriser<double> c0(F0);
    c0.factorise();
    c0.solve(X0);
    DTensor<double> A0(a, n, n, k);
    P0.addAB(A0, X0);
    Svd<double> s0(A0, true, false);
    s0.factorise();
    // the same on stream 1 while stream 0 runs an unrelated batch of products
    DTensor<float> G(std::vector<float>(g1.begin(), g1.end()), 64, 64, 512), H(64, 64, 512);
    DTensor<double> F1(a, n, n, k), X1(bv, n, 2, k), P1(n, 2, k), A1(a, n, n, k);
    F1.setStreamIdx(1);
    X1.setStreamIdx(1);
    P1.setStreamIdx(1);
    A1.setStreamIdx(1);
    CHECK(F1.streamIdx() == 1 && F1.stream() != G.stream());
    for (int rep = 0; rep < 4; rep++) H.addAB(G, G);
    CholeskyBatchFactoriser<double> c1(F1);
    c1.factorise();
    c1.solve(X1);
    P1.addAB(A1, X1);
    Svd<double> s1(A1, true, false);
    s1.factorise();
    Session::getInstance().synchronizeAllStreams();
    CHECK(host::download(F1) == host::download(F0));
    CHECK(host::download(X1) == host::download(X0));
    CHECK(host::download(P1) == host::download(P0));
    CHECK(host::download(s1.singularValues()) == host::download(s0.singularValues()));
    CHECK(host::download(c1.info()) == host::download(c0.info()));
    CHECK_THROWS(std::invalid_argument, F1.setStreamIdx(2));
}

/* =========================================================== main ===== */

int main(int argc, char **argv) {
    Session::setStreams(2);
    const std::string filter = argc > 1 ? argv[1] : "";
    int failed = 0, ran = 0;
    for (const auto &sc : suite::registry()) {
        if (!filter.empty() && sc.name.find(filter) == std::string::npos) continue;
        suite::g_failures = 0;
        std::printf("[run ] %s\n", sc.name.c_str());
        std::fflush(stdout);
        try {
            sc.body();
        } catch (const std::exception &e) {
            ++suite::g_failures;
            std::fprintf(stderr, "    unexpected exception: %s\n", e.what());
        }
        ++ran;
        if (suite::g_failures) ++failed;
        std::printf("[%s] %s\n", suite::g_failures ? "FAIL" : " ok ", sc.name.c_str());
        std::fflush(stdout);
    }
    std::printf("%d scenarios, %d failed\n", ran, failed);
    return failed;
}
