// test_api.cpp -- the reference suite (GPUtils test/testTensor.cu) ported to
// the drop-in header include/tensor.hpp, with a minimal self-contained
// harness in place of googletest.  Each TEST mirrors the reference test of
// the same name (cited as X:LINE); template helpers are instantiated for the
// same types as there.  Runs on the GPU; exit status = number of failures.
#include <cmath>
#include <cstdio>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

#include "tensor.hpp"

#define PRECISION_LOW 1e-4
#define PRECISION_HIGH 1e-10

namespace harness {
struct Case {
    const char *name;
    std::function<void()> fn;
};
std::vector<Case> &cases() {
    static std::vector<Case> c;
    return c;
}
int failures_in_case = 0;
struct Reg {
    Reg(const char *n, std::function<void()> f) { cases().push_back({n, f}); }
};
}  // namespace harness

#define TEST_F(fixture, name)                                                 \
    static void fixture##_##name();                                           \
    static harness::Reg reg_##fixture##_##name(#fixture "." #name, fixture##_##name); \
    static void fixture##_##name()

#define FAILMSG(...)                                                                       \
    do {                                                                                   \
        ++harness::failures_in_case;                                                       \
        std::fprintf(stderr, "  %s:%d: ", __FILE__, __LINE__);                             \
        std::fprintf(stderr, __VA_ARGS__);                                                 \
        std::fprintf(stderr, "\n");                                                        \
    } while (0)
#define EXPECT_TRUE(c) \
    do { if (!(c)) FAILMSG("expected true: %s", #c); } while (0)
#define ASSERT_TRUE(c) \
    do { if (!(c)) { FAILMSG("assertion: %s", #c); return; } } while (0)
#define EXPECT_EQ(a, b) \
    do { if (!((a) == (b))) FAILMSG("expected %s == %s", #a, #b); } while (0)
#define ASSERT_EQ(a, b) \
    do { if (!((a) == (b))) { FAILMSG("assertion %s == %s", #a, #b); return; } } while (0)
#define EXPECT_NE(a, b) \
    do { if ((a) == (b)) FAILMSG("expected %s != %s", #a, #b); } while (0)
#define EXPECT_NEAR(a, b, tol)                                                                      \
    do {                                                                                            \
        double a_ = (double)(a), b_ = (double)(b);                                                  \
        if (!(std::fabs(a_ - b_) <= (double)(tol))) FAILMSG("|%s - %s| = %.3e > %.1e", #a, #b,      \
                                                            std::fabs(a_ - b_), (double)(tol));     \
    } while (0)
#define EXPECT_LT(a, b) \
    do { if (!((a) < (b))) FAILMSG("expected %s < %s (%.3e vs %.3e)", #a, #b, (double)(a), (double)(b)); } while (0)
#define ASSERT_LT(a, b) \
    do { if (!((a) < (b))) { FAILMSG("assertion %s < %s", #a, #b); return; } } while (0)
#define EXPECT_THROW(stmt, exc)                                 \
    do {                                                        \
        bool thrown_ = false;                                   \
        try { stmt; } catch (const exc &) { thrown_ = true; }   \
        if (!thrown_) FAILMSG("expected %s to throw", #stmt);   \
    } while (0)

#define TENSOR_DATA_234A {1, 2, 3, 4, 5, 6, 7, 8, 9, 8, 7, 10, 5, 4, 3, 2, 1, -1, 4, 3, 4, 3, 4, 8}
#define TENSOR_DATA_234B {7, -6, 9, 2, 1, 11, 34, -1, -4, -3, 12, 7, 9, 9, 2, 9, -9, -3, 2, 5, 4, -5, 4, 5}
#define TENSOR_DATA_234APB {8, -4, 12, 6, 6, 17, 41, 7, 5, 5, 19, 17, 14, 13, 5, 11, -8, -4, 6, 8, 8, -2, 8, 13}
#define TENSOR_DATA_234AMB {-6, 8, -6, 2, 4, -5, -27, 9, 13, 11, -5, 3, -4, -5, 1, -7, 10, 2, 2, -2, 0, 8, 0, 3}

/* ======================================================= TensorTest ===== */

template <typename T>
void tensorConstructionZero() {  // X:30-46
    DTensor<T> zero(2, 3, 4, true);
    EXPECT_EQ(2u, zero.numRows());
    EXPECT_EQ(3u, zero.numCols());
    EXPECT_EQ(4u, zero.numMats());
    std::vector<T> expected(2 * 3 * 4, 0), down(2 * 3 * 4);
    zero.download(down);
    EXPECT_TRUE(expected == down);
}
TEST_F(TensorTest, tensorConstructionZero) {
    tensorConstructionZero<float>();
    tensorConstructionZero<double>();
    tensorConstructionZero<int>();
}

template <typename T>
void tensorConstructionStorageMode() {  // X:52-106
    size_t rows = 3, cols = 2, mats = 2;
    std::vector<T> aCm = {1, 3, 5, 2, 4, 6}, bCm = {7, 9, 11, 8, 10, 12};
    const std::vector<T> Cm = {1, 3, 5, 2, 4, 6, 7, 9, 11, 8, 10, 12};
    std::vector<T> aRm = {1, 2, 3, 4, 5, 6}, bRm = {7, 8, 9, 10, 11, 12};
    std::vector<T> Rm = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    std::vector<T> hostData(rows * cols * mats);
    DTensor<T> testCm(Cm, rows, cols, mats, columnMajor);
    DTensor<T> testRm(Rm, rows, cols, mats, rowMajor);
    testCm.download(hostData);
    EXPECT_TRUE(Cm == hostData);
    testRm.download(hostData);
    EXPECT_TRUE(Cm == hostData);
    DTensor<T> testSplitCm(rows, cols, mats);
    DTensor<T> ACm(testSplitCm, 2, 0, 0), BCm(testSplitCm, 2, 1, 1);
    ACm.upload(aCm, columnMajor);
    BCm.upload(bCm, columnMajor);
    DTensor<T> testSplitRm(rows, cols, mats);
    DTensor<T> ARm(testSplitRm, 2, 0, 0), BRm(testSplitRm, 2, 1, 1);
    ARm.upload(aRm, rowMajor);
    BRm.upload(bRm, rowMajor);
    testSplitCm.download(hostData);
    EXPECT_TRUE(Cm == hostData);
    testSplitRm.download(hostData);
    EXPECT_TRUE(Cm == hostData);
}
TEST_F(TensorTest, tensorConstructionStorageMode) {
    tensorConstructionStorageMode<float>();
    tensorConstructionStorageMode<double>();
    tensorConstructionStorageMode<int>();
}

template <typename T>
void randomTensorCreation() {  // X:112-127
    auto r = DTensor<T>::createRandomTensor(20, 40, 60, -1, 1);
    EXPECT_EQ(20u, r.numRows());
    EXPECT_EQ(40u, r.numCols());
    EXPECT_EQ(60u, r.numMats());
    auto rEle = r(19, 39, 59);
    EXPECT_TRUE(rEle >= -1 && rEle <= 1);
}
TEST_F(TensorTest, randomTensorCreation) {
    randomTensorCreation<float>();
    randomTensorCreation<double>();
    randomTensorCreation<int>();
}

template <typename T>
void parseTensorFromFile(const char *fName) {  // X:133-186
    for (size_t i = 0; i < 3; i++) {
        size_t nR = 20, nC = 40, nM = 6;
        auto r = DTensor<T>::createRandomTensor(nR, nC, nM, -1, 1);
        r.saveToFile(fName);
        auto a = DTensor<T>::parseFromFile(fName);
        EXPECT_EQ(nR, a.numRows());
        EXPECT_EQ(nC, a.numCols());
        EXPECT_EQ(nM, a.numMats());
        auto diff = a - r;
        T err = diff.maxAbs();
        EXPECT_LT(err, 2 * std::numeric_limits<T>::epsilon());
    }
}
TEST_F(TensorTest, parseTensorFromFile) {
    parseTensorFromFile<float>("myTest.dtensor");
    parseTensorFromFile<double>("myTest.dtensor");
}
TEST_F(TensorTest, parseTensorFromFileBinary) {
    parseTensorFromFile<float>("myTest.bt");
    parseTensorFromFile<double>("myTest.bt");
}

TEST_F(TensorTest, parseTensorFromBinaryPython) {  // X:193-202 (file written by the .bt writer in tests/golden)
    const char *env = std::getenv("GPUTILS_GOLDEN_DIR");
    std::string fName = std::string(env ? env : "tests/golden") + "/b_d.bt";
    DTensor<double> b = DTensor<double>::parseFromFile(fName);
    for (size_t i = 0; i < 3; i++)
        for (size_t j = 0; j < 3; j++) {
            EXPECT_NEAR(1 + 2 * j + 6 * i, b(i, j, 0), PRECISION_HIGH);
            EXPECT_NEAR(2 + 2 * j + 6 * i, b(i, j, 1), PRECISION_HIGH);
        }
}

TEST_F(TensorTest, parseTensorFromNonexistentFile) {  // X:209-214
    EXPECT_THROW(DTensor<double>::parseFromFile("whatever.bt", rowMajor), std::invalid_argument);
    EXPECT_THROW(DTensor<double>::parseFromFile("whatever.txt", rowMajor), std::invalid_argument);
}

template <typename T>
void tensorMoveConstructor() {  // X:221-234
    DTensor<T> zero(2, 3, 4, true);
    DTensor<T> x(std::move(zero));
    DTensor<T> y(DTensor<T>{100, 10, 1000});
    EXPECT_EQ(0u, zero.numEl());
    EXPECT_EQ(24u, x.numEl());
}
TEST_F(TensorTest, tensorMoveConstructor) {
    tensorMoveConstructor<float>();
    tensorMoveConstructor<double>();
    tensorMoveConstructor<int>();
    tensorMoveConstructor<int *>();
    tensorMoveConstructor<double *>();
}

template <typename T>
void tensorConstructionFromVector() {  // X:241-255
    std::vector<T> data = TENSOR_DATA_234A;
    DTensor<T> tenz(data, 2, 3, 4);
    EXPECT_EQ(2u, tenz.numRows());
    EXPECT_EQ(3u, tenz.numCols());
    EXPECT_EQ(4u, tenz.numMats());
    EXPECT_EQ(24u, tenz.numEl());
}
TEST_F(TensorTest, tensorConstructionFromVector) {
    tensorConstructionFromVector<float>();
    tensorConstructionFromVector<double>();
    tensorConstructionFromVector<int>();
}

template <typename T>
void tensorCopyConstructor() {  // X:261-280
    std::vector<T> data = TENSOR_DATA_234A;
    DTensor<T> tenz(data, 2, 3, 4);
    DTensor<T> tenzCp(tenz);
    EXPECT_EQ(24u, tenzCp.numEl());
    std::vector<T> down(24);
    tenzCp.download(down);
    EXPECT_TRUE(data == down);
    EXPECT_NE(tenz.raw(), tenzCp.raw());
}
TEST_F(TensorTest, tensorCopyConstructor) {
    tensorCopyConstructor<float>();
    tensorCopyConstructor<double>();
    tensorCopyConstructor<int>();
}

template <typename T>
void tensorSlicing() {  // X:287-355
    std::vector<T> data = TENSOR_DATA_234A;
    DTensor<T> tens(data, 2, 3, 4);
    DTensor<T> s2(tens, 2, 0, 1);
    EXPECT_EQ(2u, s2.numMats());
    EXPECT_EQ(tens.raw(), s2.raw());
    EXPECT_TRUE(s2.ptrMatrices() != nullptr);
    DTensor<T> s1(tens, 1, 1, 2);
    EXPECT_EQ(2u, s1.numRows());
    EXPECT_EQ(2u, s1.numCols());
    EXPECT_EQ(1u, s1.numMats());
    EXPECT_TRUE(s1.ptrMatrices() == nullptr);
    std::vector<T> e1 = {3, 4, 5, 6}, d1(4);
    s1.download(d1);
    EXPECT_TRUE(e1 == d1);
    DTensor<T> s0(tens, 0, 2, 3);
    EXPECT_EQ(2u, s0.numRows());
    EXPECT_EQ(1u, s0.numCols());
    std::vector<T> e0 = {3, 4}, d0(2);
    s0.download(d0);
    EXPECT_TRUE(e0 == d0);
    DTensor<T> s23(tens, 2, 2, 3);
    EXPECT_TRUE(s23.ptrMatrices() == tens.ptrMatrices() + 2);  // X:860-872
}
TEST_F(TensorTest, tensorSlicingConstructors) {
    tensorSlicing<float>();
    tensorSlicing<double>();
    tensorSlicing<int>();
}

template <typename T>
void tensorUploadBracketDeviceCopy() {  // X:361-404, 653-666
    std::vector<T> data = TENSOR_DATA_234A;
    DTensor<T> tenz(2, 3, 4);
    tenz.upload(data);
    EXPECT_EQ(8, tenz(1, 2, 3));
    EXPECT_EQ(1, tenz(0, 0, 0));
    EXPECT_EQ(3, tenz(0, 1, 2));
    DTensor<T> other(2, 3, 5, true);
    DTensor<T> z(other, 2, 1, 4);
    tenz.deviceCopyTo(z);
    std::vector<T> expected = {0, 0, 0, 0, 0, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 8, 7, 10, 5, 4, 3, 2, 1, -1, 4, 3, 4, 3, 4, 8};
    std::vector<T> actual(30);
    other.download(actual);
    EXPECT_TRUE(expected == actual);
}
TEST_F(TensorTest, tensorUploadBracketDeviceCopyTo) {
    tensorUploadBracketDeviceCopy<float>();
    tensorUploadBracketDeviceCopy<double>();
    tensorUploadBracketDeviceCopy<int>();
}

template <typename T>
void tensorReshape() {  // X:411-448
    size_t m = 5, n = 10, k = 3;
    DTensor<T> a = DTensor<T>::createRandomTensor(m, n, k, -1, 1);
    T lastElement = a(m - 1, n - 1, k - 1), firstElement = a(0, 0, 0);
    a.reshape(m, k, n);
    ASSERT_EQ(n, a.numMats());
    a.reshape(k, n, m);
    ASSERT_EQ(m, a.numMats());
    a.reshape(k * n, m, 1);
    ASSERT_EQ(1u, a.numMats());
    a.reshape(m, k * n, 1);
    a.reshape(m * k * n, 1, 1);
    ASSERT_EQ(m * k * n, a.numRows());
    ASSERT_EQ(lastElement, a(m * n * k - 1, 0, 0));
    ASSERT_EQ(firstElement, a(0, 0, 0));
    EXPECT_THROW(a.reshape(7, 7, 7), std::invalid_argument);
}
TEST_F(TensorTest, tensorReshape) {
    const size_t before = Session::getInstance().totalAllocatedBytes();
    tensorReshape<float>();
    tensorReshape<double>();
    tensorReshape<int>();
    EXPECT_EQ(before, Session::getInstance().totalAllocatedBytes());  // exact accounting
}

template <typename T>
void tensorSliceAndReshape(T epsilon) {  // X:454-482
    std::vector<T> dataA = TENSOR_DATA_234A, dataB = TENSOR_DATA_234B;
    DTensor<T> a(dataA, 2, 3, 4), b(dataB, 2, 3, 4);
    DTensor<T> aSlice(a, 2, 1, 3), bSlice(b, 2, 1, 3);
    aSlice.reshape(2, 9, 1);
    bSlice.reshape(2, 9, 1);
    aSlice += bSlice;
    std::vector<T> dataAExpected = {1, 2, 3, 4, 5, 6, 41, 7, 5, 5, 19, 17, 14, 13, 5, 11, -8, -4, 6, 8, 8, -2, 8, 13};
    DTensor<T> aExpected(dataAExpected, 2, 3, 4);
    DTensor<T> err = aExpected - a;
    ASSERT_LT(err.normF(), epsilon);
}
TEST_F(TensorTest, tensorSliceAndReshape) {
    tensorSliceAndReshape<float>(PRECISION_LOW);
    tensorSliceAndReshape<double>(PRECISION_HIGH);
}

template <typename T>
void tensorReductions(T epsilon) {  // X:488-573
    std::vector<T> dataA = TENSOR_DATA_234A, dataB = TENSOR_DATA_234B, dataAmB = TENSOR_DATA_234AMB;
    DTensor<T> vecA(dataA, dataA.size()), vecB(dataB, dataB.size());
    EXPECT_EQ(604, vecA.dotF(vecB));
    DTensor<T> tenA(dataA, 2, 3, 4), tenB(dataB, 2, 3, 4);
    EXPECT_EQ(604, tenA.dotF(tenB));
    EXPECT_NEAR(26.153393661244042, tenA.normF(), epsilon);
    EXPECT_NEAR(112, tenA.sumAbs(), PRECISION_HIGH);
    DTensor<T> amb(dataAmB, 2, 3, 4);
    EXPECT_EQ(27, amb.maxAbs());
    EXPECT_EQ(0, amb.minAbs());
    EXPECT_THROW(tenA.dotF(vecB), std::invalid_argument);
}
TEST_F(TensorTest, tensorReductions) {
    tensorReductions<float>(PRECISION_LOW);
    tensorReductions<double>(PRECISION_HIGH);
}

template <typename T>
void tensorAssignmentOperator() {  // X:672-688
    std::vector<T> data = TENSOR_DATA_234A;
    DTensor<T> tenz(data, 2, 3, 4);
    DTensor<T> other;
    other = tenz;
    EXPECT_EQ(tenz.raw(), other.raw());
    EXPECT_EQ(4u, other.numMats());
}
TEST_F(TensorTest, tensorAssignmentOperator) {
    tensorAssignmentOperator<float>();
    tensorAssignmentOperator<double>();
    tensorAssignmentOperator<int>();
}

template <typename T>
void tensorArithmetic() {  // X:694-822
    std::vector<T> data = TENSOR_DATA_234A, dataB = TENSOR_DATA_234B;
    std::vector<T> times3 = {3, 6, 9, 12, 15, 18, 21, 24, 27, 24, 21, 30, 15, 12, 9, 6, 3, -3, 12, 9, 12, 9, 12, 24};
    std::vector<T> apb = TENSOR_DATA_234APB, amb = TENSOR_DATA_234AMB, actual;
    DTensor<T> t1(data, 2, 3, 4);
    t1 *= 3.0;
    t1.download(actual);
    EXPECT_TRUE(times3 == actual);
    DTensor<T> t2(data, 2, 3, 4);
    auto t3 = (T)3.0 * t2;
    t3.download(actual);
    EXPECT_TRUE(times3 == actual);
    DTensor<T> A(data, 2, 3, 4), B(dataB, 2, 3, 4);
    A += B;
    A.download(actual);
    EXPECT_TRUE(apb == actual);
    DTensor<T> A2(data, 2, 3, 4);
    A2 -= B;
    A2.download(actual);
    EXPECT_TRUE(amb == actual);
    DTensor<T> A3(data, 2, 3, 4);
    DTensor<T> C = A3 + B;
    C.download(actual);
    EXPECT_TRUE(apb == actual);
    DTensor<T> D = A3 - B;
    D.download(actual);
    EXPECT_TRUE(amb == actual);
}
TEST_F(TensorTest, tensorArithmetic) {
    tensorArithmetic<float>();
    tensorArithmetic<double>();
}

template <typename T>
void tensorAddAB() {  // X:828-853
    std::vector<T> aData = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18};
    std::vector<T> bData = {6, 5, 4, 3, 2, 1, 7, 6, 5, 4, 3, 2, 1, 2, 1, 5, -6, 8};
    DTensor<T> A(aData, 2, 3, 3), B(bData, 3, 2, 3);
    DTensor<T> C(2, 2, 3, true);
    C.addAB(A, B);
    std::vector<T> expected = {41, 56, 14, 20, 158, 176, 77, 86, 60, 64, 111, 118}, actual;
    C.download(actual);
    EXPECT_TRUE(expected == actual);
}
TEST_F(TensorTest, tensorAddAB) {
    tensorAddAB<double>();
    tensorAddAB<float>();
}

template <typename T>
void tensorGetRows() {  // X:899-926
    std::vector<T> aData = {10.5, 25.0, 60.0, -21.0, 720.0, -1.0, 11.0, -1.0, 30.0, 5., 6., 7., 8., 9., 10., 11., 12., 13};
    DTensor<T> A(aData, 3, 3, 2);
    DTensor<T> Ar0 = A.getRows(1, 1, 0);
    std::vector<T> expected0 = {25., 720., -1.}, actual0(3);
    Ar0.download(actual0);
    EXPECT_TRUE(expected0 == actual0);
    DTensor<T> Ar1 = A.getRows(1, 2, 1);
    std::vector<T> expected1 = {6., 7., 9., 10., 12., 13.}, actual1(6);
    Ar1.download(actual1);
    EXPECT_TRUE(expected1 == actual1);
}
TEST_F(TensorTest, tensorGetRows) {
    tensorGetRows<float>();
    tensorGetRows<double>();
}

template <typename T>
void tensorTranspose() {  // X:933-950
    std::vector<T> aData = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    DTensor<T> A(aData, 3, 2, 2);
    DTensor<T> At = A.tr();
    EXPECT_EQ(2u, At.numRows());
    EXPECT_EQ(3u, At.numCols());
    std::vector<T> expected = {1, 4, 2, 5, 3, 6, 7, 10, 8, 11, 9, 12}, actual;
    At.download(actual);
    EXPECT_TRUE(expected == actual);
}
TEST_F(TensorTest, tensorTranspose) {
    tensorTranspose<float>();
    tensorTranspose<double>();
}

template <typename T>
void tensorBytesAllocated() {  // X:955-992
    const size_t before = Session::getInstance().totalAllocatedBytes();
    size_t m = 10, n = 10, k = 20;
    auto *A = new DTensor<T>(m, n, k);
    EXPECT_EQ(before + m * n * k * sizeof(T) + k * sizeof(T *), Session::getInstance().totalAllocatedBytes());
    delete A;
    EXPECT_EQ(before, Session::getInstance().totalAllocatedBytes());
}
TEST_F(TensorTest, tensorBytesAllocated) {
    tensorBytesAllocated<float>();
    tensorBytesAllocated<double>();
    tensorBytesAllocated<int>();
}

/* ===================================================== CholeskyTest ===== */

template <typename T>
void choleskyFactorisation(T epsilon) {  // X:1218-1238
    std::vector<T> aData = {10.0, 2.0, 3.0, 2.0, 20.0, -1.0, 3.0, -1.0, 30.0};
    DTensor<T> A(aData, 3, 3, 1);
    CholeskyFactoriser<T> chol(A);
    chol.factorise();
    EXPECT_NEAR(3.162277660168380, A(0, 0), epsilon);
    EXPECT_NEAR(-0.361403161162101, A(2, 1), epsilon);
    EXPECT_NEAR(5.382321781081287, A(2, 2), epsilon);
    EXPECT_EQ(0, chol.info()(0));
}
TEST_F(CholeskyTest, choleskyFactorisation) {
    choleskyFactorisation<float>(PRECISION_LOW);
    choleskyFactorisation<double>(PRECISION_HIGH);
}

template <typename T>
void choleskyFactorisationSolution(T epsilon) {  // X:1244-1276
    std::vector<T> aData = {10.0, 2.0, 3.0, 2.0, 20.0, -1.0, 3.0, -1.0, 30.0};
    DTensor<T> A(aData, 3, 3, 1);
    DTensor<T> L(A);
    CholeskyFactoriser<T> chol(L);
    chol.factorise();
    std::vector<T> bData = {-1., -3., 5.};
    DTensor<T> rhs(bData, 3, 1, 1);
    DTensor<T> sol(rhs);
    chol.solve(sol);
    std::vector<T> expected = {-0.126805213103205, -0.128566396618528, 0.175061641423036}, actual(3);
    sol.download(actual);
    for (size_t i = 0; i < 3; i++) EXPECT_NEAR(expected[i], actual[i], epsilon);
    DTensor<T> error = A * sol;
    error -= rhs;
    EXPECT_TRUE(error.normF() < epsilon);
}
TEST_F(CholeskyTest, choleskyFactorisationSolution) {
    choleskyFactorisationSolution<float>(PRECISION_LOW);
    choleskyFactorisationSolution<double>(PRECISION_HIGH);
}

template <typename T>
void choleskyBatch(T epsilon, bool provided) {  // X:1282-1413
    std::vector<T> aData = {10.0, 2.0, 3.0, 2.0, 20.0, -1.0, 3.0, -1.0, 30.0};
    DTensor<T> A(3, 3, 2);
    DTensor<T> A0(A, 2, 0, 0), A1(A, 2, 1, 1);
    A0.upload(aData);
    A1.upload(aData);
    std::vector<T> lowData = {3.162277660168380, 0, 0, 0.632455532033676, 4.427188724235731, 0,
                              0.948683298050514, -0.361403161162101, 5.382321781081287};
    DTensor<T> low(3, 3, 2);
    DTensor<T> low0(low, 2, 0, 0), low1(low, 2, 1, 1);
    low0.upload(lowData, rowMajor);
    low1.upload(lowData, rowMajor);
    DTensor<T> L(provided ? low : A);
    CholeskyBatchFactoriser<T> chol(L, provided);
    chol.factorise();  // no-op when the factor is provided
    if (!provided) {
        EXPECT_NEAR(3.162277660168380, L(0, 0, 1), epsilon);
        EXPECT_NEAR(-0.361403161162101, L(2, 1, 1), epsilon);
        EXPECT_NEAR(5.382321781081287, L(2, 2, 0), epsilon);
    }
    std::vector<T> bData = {-1., -3., 5.};
    DTensor<T> rhs(3, 1, 2);
    DTensor<T> rhs0(rhs, 2, 0, 0), rhs1(rhs, 2, 1, 1);
    rhs0.upload(bData);
    rhs1.upload(bData);
    DTensor<T> sol(rhs);
    chol.solve(sol);
    std::vector<T> expected = {-0.126805213103205, -0.128566396618528, 0.175061641423036}, actual(6);
    sol.download(actual);
    for (size_t i = 0; i < 3; i++) {
        EXPECT_NEAR(expected[i], actual[i], epsilon);
        EXPECT_NEAR(expected[i], actual[i + 3], epsilon);
    }
    DTensor<T> error = A * sol;
    error -= rhs;
    EXPECT_TRUE(error.normF() < epsilon);
    EXPECT_EQ(0, chol.info()(0));
}
TEST_F(CholeskyTest, choleskyBatchFactorSolve) {
    choleskyBatch<float>(PRECISION_LOW, false);
    choleskyBatch<double>(PRECISION_HIGH, false);
}
TEST_F(CholeskyTest, choleskyBatchSolve) {
    choleskyBatch<float>(PRECISION_LOW, true);
    choleskyBatch<double>(PRECISION_HIGH, true);
}

TEST_F(CholeskyTest, choleskyBatchPreconditions) {  // T:2117, 2163-2167
    DTensor<double> notSquare(3, 2, 2);
    EXPECT_THROW(CholeskyBatchFactoriser<double> c(notSquare), std::invalid_argument);
    DTensor<double> A(3, 3, 2, true);
    CholeskyBatchFactoriser<double> chol(A);
    DTensor<double> b(3, 1, 2, true);
    EXPECT_THROW(chol.solve(b), std::logic_error);
    CholeskyBatchFactoriser<double> chol2(A, true);
    DTensor<double> bad(4, 1, 2, true);
    EXPECT_THROW(chol2.solve(bad), std::invalid_argument);
}

TEST_F(CholeskyTest, choleskyBatchNotPositiveDefinite) {  // info convention (LAPACK)
    std::vector<double> aData = {1, 2, 2, 1};  // leading minor of order 2 is indefinite
    DTensor<double> A(aData, 2, 2, 1);
    CholeskyFactoriser<double> chol(A);
    chol.factorise();
    EXPECT_EQ(2, chol.info()(0));
}

/* ============================================================ main ====== */
int main(int argc, char **argv) {
    const char *filter = argc > 1 ? argv[1] : nullptr;
    int failed_cases = 0, run = 0;
    for (auto &c : harness::cases()) {
        if (filter && std::string(c.name).find(filter) == std::string::npos) continue;
        harness::failures_in_case = 0;
        std::printf("[ RUN      ] %s\n", c.name);
        std::fflush(stdout);
        try {
            c.fn();
        } catch (const std::exception &e) {
            ++harness::failures_in_case;
            std::fprintf(stderr, "  exception: %s\n", e.what());
        }
        ++run;
        if (harness::failures_in_case) {
            ++failed_cases;
            std::printf("[  FAILED  ] %s\n", c.name);
        } else {
            std::printf("[       OK ] %s\n", c.name);
        }
        std::fflush(stdout);
    }
    std::printf("%d tests, %d failed\n", run, failed_cases);
    return failed_cases;
}
