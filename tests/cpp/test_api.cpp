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

/* ================================================================ LS ====== */

template <typename T>
void tensorLeastSquares1(T epsilon) {  // X:1012-1034
    std::vector<T> aData = {1, 2, 3, 4, 7, 8, 9, 10, 6, 8, -9, 20};
    std::vector<T> bData = {1, 1, -1, 2, 30, -80};
    DTensor<T> A0(aData, 2, 2, 3);
    DTensor<T> A(A0);
    DTensor<T> B(bData, 2, 1, 3);
    DTensor<T> sol(B);
    A0.leastSquaresBatched(sol);
    DTensor<T> C(2, 1, 3);
    C.addAB(A, sol);
    C -= B;
    EXPECT_LT(C.normF(), epsilon);
}

TEST_F(LeastSquaresTest, tensorLS1) {
    tensorLeastSquares1<float>(PRECISION_LOW);
    tensorLeastSquares1<double>(PRECISION_HIGH);
}

TEST_F(LeastSquaresTest, tensorLSPreconditions) {  // T:1344-1351
    DTensor<double> A(3, 2, 2), fat(2, 3, 2), b(3, 1, 2), b2(3, 2, 2), bk(3, 1, 1), b4(4, 1, 2);
    EXPECT_THROW(A.leastSquaresBatched(b4), std::invalid_argument);
    EXPECT_THROW(A.leastSquaresBatched(b2), std::invalid_argument);
    EXPECT_THROW(A.leastSquaresBatched(bk), std::invalid_argument);
    DTensor<double> bf(2, 1, 2);
    EXPECT_THROW(fat.leastSquaresBatched(bf), std::invalid_argument);
}

/* ================================================================ QR ====== */

template <typename T>
void qrFactorisationOfSize(size_t nR, size_t nC, T epsilon) {  // X:1433-1490
    DTensor<T> temp(nR, nC);
    DTensor<T> A = DTensor<T>::createRandomTensor(nR, nC, 1, -100, 100);
    QRFactoriser<T> qr(temp);
    A.deviceCopyTo(temp);
    qr.factorise();
    EXPECT_EQ(0, qr.info()(0));
    DTensor<T> Q(nR, nC);
    DTensor<T> R(nC, nC, 1, true);
    DTensor<T> QR(nR, nC);
    qr.getQR(Q, R);
    EXPECT_EQ(0, qr.info()(0));
    QR.addAB(Q, R);
    QR -= A;
    EXPECT_NEAR(QR.normF(), 0., epsilon);
}

TEST_F(QRTest, qrFactorisation) {
    qrFactorisationOfSize<float>(4, 3, PRECISION_LOW);
    qrFactorisationOfSize<double>(4, 3, PRECISION_HIGH);
}

TEST_F(QRTest, qrFactorisationTall) {
    qrFactorisationOfSize<float>(20, 3, PRECISION_LOW);
    qrFactorisationOfSize<double>(20, 3, PRECISION_HIGH);
}

template <typename T>
void qrLeastSquares(T epsilon) {  // X:1496-1530
    const size_t nR = 4, nC = 3;
    DTensor<T> temp(nR, nC);
    std::vector<T> vecA = {85.5638, -59.4001, -80.1992, 99.9464, 5.51393, 5.17935,
                           6.87488, -26.7536, 36.0914, -44.3857, -32.1268, 54.8915};
    std::vector<T> vecB = {-23.3585, -48.5744, 43.4229, -56.5081};
    DTensor<T> A(vecA, nR, nC, 1, rowMajor);
    DTensor<T> b(vecB, nR);
    DTensor<T> xFull(nR);
    DTensor<T> x(xFull, 0, 0, nC - 1);
    DTensor<T> Ax(nR);
    QRFactoriser<T> qr(temp);
    A.deviceCopyTo(temp);
    qr.factorise();
    EXPECT_EQ(0, qr.info()(0));
    b.deviceCopyTo(xFull);
    qr.leastSquares(xFull);
    EXPECT_EQ(0, qr.info()(0));
    Ax.addAB(A, x);
    Ax -= b;
    EXPECT_NEAR(Ax.normF(), 80.003169364198072, epsilon);
}

TEST_F(QRTest, qrLeastSquares) {
    qrLeastSquares<float>(PRECISION_LOW);
    qrLeastSquares<double>(PRECISION_HIGH);
}

TEST_F(QRTest, qrPreconditions) {  // T:1811-1813, 1933-1936
    DTensor<double> fat(2, 3), batch(3, 2, 2), A(4, 3), Qbad(4, 2), R(3, 3);
    EXPECT_THROW(QRFactoriser<double> q(fat), std::invalid_argument);
    EXPECT_THROW(QRFactoriser<double> q(batch), std::invalid_argument);
    QRFactoriser<double> qr(A);
    EXPECT_THROW(qr.getQR(Qbad, R), std::invalid_argument);
}

/* =============================================================== SVD ====== */

template <typename T>
void singularValuesComputation(T epsilon) {  // X:1058-1078
    std::vector<T> bData = {1, 6, 6, 6, 6, 6, 6, 6, 2, 7, 7, 7, 7, 7, 7, 7, 3, 8, 8, 8, 8, 8, 8, 8};
    DTensor<T> B(bData, 8, 3);
    Svd<T> svd(B, true, false);
    svd.factorise();
    for (size_t i = 0; i < B.numMats(); i++) EXPECT_EQ(0, svd.info()(i));
    auto S = svd.singularValues();
    EXPECT_NEAR(32.496241123753592, S(0), epsilon);
    EXPECT_NEAR(0.997152358903242, S(1), epsilon);
    auto U = svd.leftSingularVectors();
    EXPECT_TRUE(U.has_value());
}

TEST_F(SvdTest, singularValuesComputation) {
    singularValuesComputation<float>(PRECISION_LOW);
    singularValuesComputation<double>(PRECISION_HIGH);
}

template <typename T>
void singularValuesMemory() {  // X:1086-1108
    std::vector<T> bData = {1, 6, 6, 6, 6, 6, 6, 6, 2, 7, 7, 7, 7, 7, 7, 7, 3, 8, 8, 8, 8, 8, 8, 8};
    DTensor<T> B(bData, 8, 3);
    Svd<T> svd(B, true, false);
    svd.factorise();
    EXPECT_EQ(0, svd.info()(0));
    DTensor<T> const &v1 = svd.rightSingularVectors();
    DTensor<T> const &v2 = svd.rightSingularVectors();
    EXPECT_EQ(&v1, &v2);
    EXPECT_EQ(v1.raw(), v2.raw());
    DTensor<T> const &s1 = svd.singularValues();
    DTensor<T> const &s2 = svd.singularValues();
    EXPECT_EQ(&s1, &s2);
    EXPECT_EQ(s1.raw(), s2.raw());
    auto u1 = svd.leftSingularVectors().value();
    auto u2 = svd.leftSingularVectors().value();
    EXPECT_EQ(u1, u2);
    EXPECT_EQ(u1->raw(), u2->raw());
}

TEST_F(SvdTest, singularValuesMemory) {
    singularValuesMemory<float>();
    singularValuesMemory<double>();
}

template <typename T>
void singularValuesMultipleMatrices(T epsilon) {  // X:1117-1160
    std::vector<T> aData = {1, 2, 3, 4, 5, 6, 1, 1, 1, 2, 2, 2, 0, 0, 0, 0, 0, 1};
    DTensor<T> A(aData, 3, 2, 3);
    Svd<T> svd(A, true);
    svd.factorise();
    DTensor<T> const &S = svd.singularValues();
    DTensor<T> const &V = svd.rightSingularVectors();
    auto U = svd.leftSingularVectors().value();
    std::vector<T> expected_v = {-0.386317703118612, -0.922365780077058, -0.922365780077058, 0.386317703118612};
    std::vector<T> actual_v(12);
    V.download(actual_v);
    for (size_t i = 0; i < 4; i++) EXPECT_NEAR(expected_v[i], actual_v[i], epsilon);
    std::vector<T> expected_s = {9.508032000695726, 0.772869635673484, 3.872983346207417, 0, 1, 0};
    std::vector<T> actual_s(6);
    S.download(actual_s);
    for (size_t i = 0; i < 6; i++) EXPECT_NEAR(expected_s[i], actual_s[i], epsilon);
    std::vector<T> expected_u = {
        -0.428667133548626, -0.566306918848035, -0.703946704147444, 0.805963908589298, 0.112382414096594,
        -0.581199080396110, 0.408248290463863,  -0.816496580927726, 0.408248290463863, -0.577350269189626,
        -0.577350269189626, -0.577350269189626, 0.816496580927726,  -0.408248290463863, -0.408248290463863,
        0.000000000000000,  -0.707106781186548, 0.707106781186547,  0, 0, -1, 1, 0, 0, 0, -1, 0};
    std::vector<T> actual_u(27);
    U->download(actual_u);
    for (size_t i = 0; i < 27; i++) EXPECT_NEAR(expected_u[i], actual_u[i], epsilon);
}

TEST_F(SvdTest, singularValuesMultipleMatrices) {
    singularValuesMultipleMatrices<float>(10 * PRECISION_LOW);
    singularValuesMultipleMatrices<double>(PRECISION_HIGH);
}

template <typename T>
void singularValuesRankMultipleMatrices(T epsilon) {  // X:1175-1188
    std::vector<T> aData = {1, 4, 7, 10, 2, 5, 8, 11, 3, 6, 9, 0, 1, 4, 7, 10, 2, 5,
                            8, 11, 3, 6, 9, 12, 1, 2, 3, 4, 2, 4, 6, 8, 3, 6, 9, 12};
    DTensor<T> A(aData, 4, 3, 3);
    Svd<T> svd(A);
    svd.factorise();
    auto rank = svd.rank(epsilon);
    EXPECT_EQ(3u, rank(0, 0, 0));
    EXPECT_EQ(2u, rank(0, 0, 1));
    EXPECT_EQ(1u, rank(0, 0, 2));
    // the count is recomputed, not accumulated (T:1602-1607 adds to it)
    auto again = svd.rank(epsilon);
    EXPECT_EQ(3u, again(0, 0, 0));
}

TEST_F(SvdTest, singularValuesRankMultipleMatrices) {
    singularValuesRankMultipleMatrices<float>(PRECISION_LOW);
    singularValuesRankMultipleMatrices<double>(PRECISION_HIGH);
}

TEST_F(SvdTest, svdPreconditions) {  // T:1518-1522
    DTensor<double> fat(2, 3);
    EXPECT_THROW(Svd<double> s(fat), std::invalid_argument);
}

/* ========================================================= Nullspace ====== */

template <typename T>
void computeNullspaceTensor(T epsilon) {  // X:1556-1585
    std::vector<T> aData = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 0, 1, 2, 3, 4, 5, 6, 7, 8,
                            9, 7, 8, 9, 1, 2, 3, 4, 2, 4, 6, 8, 3, 6, 9, 12, 1, 1, 1, 1,
                            1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    DTensor<T> A(aData, 3, 4, 5);
    Nullspace<T> ns(A);
    DTensor<T> nA = ns.nullspace();
    const size_t nMats = nA.numMats();
    EXPECT_EQ(nMats, 5u);
    for (size_t i = 0; i < nMats; i++) {
        DTensor<T> nAi(nA, 2, i, i);
        DTensor<T> Ai(A, 2, i, i);
        DTensor<T> mustBeZero = Ai * nAi;
        EXPECT_LT(mustBeZero.normF(), epsilon);
        DTensor<T> nAiTr = nAi.tr();
        DTensor<T> mustBeEye = nAiTr * nAi;
        EXPECT_NEAR(1, mustBeEye(0, 0, 0), epsilon);
        for (size_t ir = 0; ir < mustBeEye.numRows(); ir++)
            for (size_t ic = 0; ic < mustBeEye.numCols(); ic++)
                if (ir != ic) EXPECT_NEAR(0, mustBeEye(ir, ic, 0), epsilon);
    }
}

TEST_F(NullspaceTest, computeNullspaceTensor) {
    computeNullspaceTensor<float>(PRECISION_LOW);
    computeNullspaceTensor<double>(PRECISION_HIGH);
}

template <typename T>
void computeNullspaceTrivial() {  // X:1597-1613
    std::vector<T> data{4, 5, 7, 4, 1, 8, 4, 5, 0, 1, 1, 1, 5, 6, 7, 9, 0, 3};
    DTensor<T> A(data, 3, 3, 2, rowMajor);
    Nullspace<T> nullA(A);
    DTensor<T> N = nullA.nullspace();
    EXPECT_EQ(N.normF(), 0);
}

TEST_F(NullspaceTest, computeNullspaceTrivial) {
    computeNullspaceTrivial<float>();
    computeNullspaceTrivial<double>();
}

template <typename T>
void projectOnNullspaceTensor(T epsilon) {  // X:1624-1657
    const size_t m = 3, n = 7;
    std::vector<T> mat{1, -2, 3, 4, -1, -1, -1, 1, 2, -3, 4, -1, -1, -1, -1, 3, 5, -7, -1, -1, -1};
    DTensor<T> A(m, n, 1);
    A.upload(mat, rowMajor);
    Nullspace<T> ns = Nullspace(A);
    DTensor<T> N = ns.nullspace();
    std::vector<T> vec{1, 2, 3, 4, 5, 6, 7};
    DTensor<T> x(vec, n);
    DTensor<T> proj(x);
    ns.project(proj);
    DTensor<T> error(m, 1, 1, true);
    error.addAB(A, proj);
    EXPECT_TRUE(error.normF() < epsilon);
    std::vector<T> h_other{1, -2, 5, 4, 0, 0, 0};
    DTensor<T> other(h_other, n);
    DTensor<T> y = N * other;
    DTensor<T> delta1 = y - proj;
    DTensor<T> delta2 = proj - x;
    EXPECT_LT(delta1.dotF(delta2), epsilon);
}

TEST_F(NullspaceTest, projectOnNullspaceTensor) {
    projectOnNullspaceTensor<float>(PRECISION_LOW);
    projectOnNullspaceTensor<double>(PRECISION_HIGH);
}

TEST_F(NullspaceTest, projectCapturedInGraph) {  // repeated project through StreamGraph
    const size_t m = 3, n = 7;
    std::vector<double> mat{1, -2, 3, 4, -1, -1, -1, 1, 2, -3, 4, -1, -1, -1, -1, 3, 5, -7, -1, -1, -1};
    DTensor<double> A(m, n, 1);
    A.upload(mat, rowMajor);
    Nullspace<double> ns(A);
    std::vector<double> vec{1, 2, 3, 4, 5, 6, 7};
    DTensor<double> eager(vec, n), b(vec, n);
    ns.project(eager);
    StreamGraph g(b.streamIdx());
    g.capture([&] { ns.project(b); });
    for (int rep = 0; rep < 3; ++rep) {
        b.upload(vec);
        g.launch();
        Session::getInstance().synchronizeAllStreams();
        std::vector<double> got(n), want(n);
        b.download(got);
        eager.download(want);
        for (size_t i = 0; i < n; i++) EXPECT_EQ(got[i], want[i]);
    }
}

/* ============================================================ Givens ==== */

// X:580-603: columns 1 and 4 of the 10 x 6 iota matrix rotated
template <typename T>
static void tensorRightGivens(T epsilon) {
    const size_t m = 10, n = 6;
    std::vector<T> v(m * n);
    std::iota(v.begin(), v.end(), T(1));
    DTensor<T> a(v, m, n, 1);
    const size_t ig = 1, jg = 4;
    const T c = 0.1, minus_s = std::sqrt(1 - c * c);
    a.applyRightGivensRotation(ig, jg, &c, &minus_s);
    for (size_t i = 0; i < m; i++) {
        EXPECT_NEAR(1 + i, a(i, 0), epsilon);
        EXPECT_NEAR(21 + i, a(i, 2), epsilon);
        EXPECT_NEAR(31 + i, a(i, 3), epsilon);
        EXPECT_NEAR((11 + i) * c - (41 + i) * (-minus_s), a(i, ig), epsilon);
        EXPECT_NEAR((11 + i) * (-minus_s) + (41 + i) * c, a(i, jg), epsilon);
    }
}

TEST_F(TensorTest, tensorRightGivens) {
    tensorRightGivens<float>(PRECISION_LOW);
    tensorRightGivens<double>(PRECISION_HIGH);
}

// X:615-641: rows 1 and 9 rotated (the reference builds a double tensor in
// both instantiations)
template <typename T>
static void tensorLeftGivens(T epsilon) {
    const size_t m = 10, n = 6;
    std::vector<double> v(m * n);
    std::iota(v.begin(), v.end(), 1.0);
    DTensor<double> a(v, m, n, 1);
    const size_t ig = 1, jg = 9;
    const double c = 0.1, minus_s = -std::sqrt(1 - c * c);
    a.applyLeftGivensRotation(ig, jg, &c, &minus_s);
    for (size_t j = 0; j < n; j++) {
        EXPECT_NEAR(1 + 10 * j, a(0, j), epsilon);
        for (size_t i = 2; i < m - 1; i++) EXPECT_NEAR(1 + i + 10 * j, a(i, j), epsilon);
        EXPECT_NEAR((2 + 10 * j) * c + (10 + 10 * j) * minus_s, a(ig, j), epsilon);
        EXPECT_NEAR((2 + 10 * j) * (-minus_s) + (10 + 10 * j) * c, a(jg, j), epsilon);
    }
}

TEST_F(TensorTest, tensorLeftGivens) {
    tensorLeftGivens<float>(1e-10);
    tensorLeftGivens<double>(1e-14);
}

// X:1684-1701: every element below row 0 annihilated in turn
template <typename T>
static void givensAnnihilateElement(T epsilon) {
    const size_t m = 10, n = 6;
    std::vector<T> v(m * n);
    std::iota(v.begin(), v.end(), T(1));
    DTensor<T> a(v, m, n, 1);
    GivensAnnihilator<T> ga(a);
    for (size_t k = 1; k < m; k++) {
        for (size_t j = 0; j < n; j++) {
            ga.annihilate(0, k, j);
            EXPECT_NEAR(0.0, a(k, j), epsilon);
        }
    }
}

TEST_F(GivensAnnihilatorTest, givensAnnihilateElement) {
    givensAnnihilateElement<float>(PRECISION_LOW);
    givensAnnihilateElement<double>(PRECISION_HIGH);
}

// X:1713-1734: golden entries after annihilating (1, 2) with row 0
template <typename T>
static void givensAnnihilateCorrectness(T epsilon) {
    const size_t m = 10, n = 6;
    std::vector<double> v(m * n);
    std::iota(v.begin(), v.end(), 1.0);
    DTensor<double> a(v, m, n);
    GivensAnnihilator<double> ga(a);
    ga.annihilate(0, 1, 2);
    EXPECT_NEAR(0.0, a(1, 2), epsilon);
    EXPECT_NEAR(2.137186834969645, a(0, 0), epsilon);
    EXPECT_NEAR(44.552125559751822, a(0, 3), epsilon);
    EXPECT_NEAR(-0.328797974610715, a(1, 3), epsilon);
}

TEST_F(GivensAnnihilatorTest, givensAnnihilateCorrectness) {
    givensAnnihilateCorrectness<double>(1e-14);
    givensAnnihilateCorrectness<float>(1e-12);
}

TEST_F(GivensAnnihilatorTest, invalidIndicesThrow) {  // no reference counterpart: j is checked too
    std::vector<double> v(12, 1.0);
    DTensor<double> a(v, 4, 3);
    GivensAnnihilator<double> ga(a);
    bool threw_row = false, threw_col = false;
    try { ga.annihilate(0, 4, 0); } catch (const std::invalid_argument &) { threw_row = true; }
    try { ga.annihilate(0, 1, 3); } catch (const std::invalid_argument &) { threw_col = true; }
    EXPECT_TRUE(threw_row);
    EXPECT_TRUE(threw_col);
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
