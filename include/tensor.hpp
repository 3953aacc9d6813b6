/*
 * tensor.hpp -- header-only C++ API of GPUtils' batched dense linear
 * algebra (DTensor, Session, IStatus, CholeskyBatchFactoriser, ...) on
 * MI355X.  Source-compatible replacement of GPUtils' include/tensor.cuh:
 * user code that includes this header instead compiles unchanged with any
 * C++17 compiler (no HIP/CUDA compiler needed) and links
 * libgputils_hip.so, whose C ABI (gputils_hip.h) runs the gfx950 kernels.
 *
 * Deliberate differences from tensor.cuh (each one documented at the
 * method):
 *  - Session exposes HIP-native names (stream(), allocate()) instead of the
 *    vendor-named cuBlasHandle/cuSolverHandle/cudaAllocate.
 *  - Solver calls run on the tensor's stream instead of always stream 0.
 *  - CholeskyBatchFactoriser::solve accepts b with any number of columns.
 *  - Svd::rank() recomputes the rank on each call instead of adding to a
 *    counter that is never reset.
 *  - Svd handles at most 64 columns and never modifies the factorised
 *    matrix (cuSOLVER's gesvd overwrites it).
 *  - reshape() keeps the allocated-bytes counter exact.
 *  - StreamGraph (new) captures repeated calls into a HIP graph.
 */
#ifndef GPUTILS_TENSOR_HPP
#define GPUTILS_TENSOR_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <memory>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gputils_hip.h"

#define DEFAULT_FPX double
#define THREADS_PER_BLOCK 512
#define TEMPLATE_WITH_TYPE_T template <typename T = DEFAULT_FPX>
#define TEMPLATE_CONSTRAINT_REQUIRES_FPX

/* ------------------------------------------------------------ helpers --- */

static std::random_device RND_DEVICE;

/** Vector of n uniform reals in [low, hi] (tensor.cuh:47-58). */
TEMPLATE_WITH_TYPE_T
std::vector<T> generateRealRandomVector(size_t n, T low, T hi) {
    std::mt19937_64 engine(RND_DEVICE());
    std::uniform_real_distribution<T> dist(low, hi);
    std::vector<T> vec(n);
    for (auto &v : vec) v = dist(engine);
    return vec;
}

/** Vector of n uniform integers in [low, hi] (tensor.cuh:60-69). */
inline std::vector<int> generateIntRandomVector(size_t n, int low, int hi) {
    std::mt19937_64 engine(RND_DEVICE());
    std::uniform_int_distribution<int> dist(low, hi);
    std::vector<int> vec(n);
    for (auto &v : vec) v = dist(engine);
    return vec;
}

/** Ceiling division (tensor.cuh:79-81). */
constexpr size_t numBlocks(size_t n, size_t threads_per_block = THREADS_PER_BLOCK) {
    return (n / threads_per_block + (n % threads_per_block != 0));
}

/** Abort with a message on a failing C-ABI status (tensor.cuh:87-113). */
#define gpuErrChk(status) \
    { gpuAssert((status), __FILE__, __LINE__); }

inline void gpuAssert(int code, const char *file, int line, bool abort = true) {
    if (code != GPUTILS_OK) {
        std::cerr << "gputils error " << code << ": " << gputils_last_error() << ", file: " << file
                  << ", line: " << line << "\n";
        if (abort) exit(code > 0 ? code : 1);
    }
}

/* ------------------------------------------------------------ session --- */

static size_t s_numStreams = 1;

/**
 * Process-wide HIP state (tensor.cuh:133-247): one blocking stream per
 * stream index, a device-byte counter and a small reduction workspace per
 * stream.
 */
class Session {
public:
    /** Number of streams; call before the first getInstance(). */
    static void setStreams(size_t numStreams) { s_numStreams = numStreams; }

    static Session &getInstance() {
        static Session instance(s_numStreams);
        return instance;
    }

    Session(Session const &) = delete;
    void operator=(Session const &) = delete;

    /** HIP stream of index idx (as void*, a hipStream_t). */
    void *stream(size_t idx = 0) const {
        if (idx >= m_streams.size()) throw std::invalid_argument("stream index out of range");
        return m_streams[idx];
    }

    /** Counted device allocation (tensor.cuh:209-215, HIP-native name). */
    int allocate(void **d, size_t s) {
        int err = gputils_malloc(d, s);
        if (err == GPUTILS_OK) m_bytesAllocated += s;
        return err;
    }

    size_t totalAllocatedBytes() const { return m_bytesAllocated; }

    /** Adjust the byte counter (can be negative). */
    void incrementAllocatedBytes(long long s) { m_bytesAllocated += s; }

    void synchronizeStream(size_t idx = 0) const {
        if (idx >= m_streams.size()) throw std::runtime_error("stream index out of range");
        gpuErrChk(gputils_stream_synchronize(m_streams[idx]));
    }

    void synchronizeAllStreams() const {
        for (size_t i = 0; i < m_streams.size(); i++) synchronizeStream(i);
    }

    /** Reduction workspace of stream idx (not counted as tensor bytes). */
    void *reduceWorkspace(size_t idx) { return m_ws[idx]; }
    /** 16-byte device scalar slot of stream idx. */
    void *scalarSlot(size_t idx) { return m_scalar[idx]; }

private:
    explicit Session(size_t numStreams) {
        if (numStreams == 0) numStreams = 1;
        m_streams.resize(numStreams);
        m_ws.resize(numStreams);
        m_scalar.resize(numStreams);
        for (size_t i = 0; i < numStreams; i++) {
            gpuErrChk(gputils_stream_create(&m_streams[i]));
            gpuErrChk(gputils_malloc(&m_ws[i], gputils_reduce_workspace_bytes()));
            gpuErrChk(gputils_malloc(&m_scalar[i], 16));
        }
    }
    ~Session() {
        for (size_t i = 0; i < m_streams.size(); i++) {
            gputils_stream_synchronize(m_streams[i]);
            gputils_free(m_ws[i]);
            gputils_free(m_scalar[i]);
            gputils_stream_destroy(m_streams[i]);
        }
    }

    std::vector<void *> m_streams;
    std::vector<void *> m_ws;
    std::vector<void *> m_scalar;
    size_t m_bytesAllocated = 0;
};

/* ------------------------------------------------------------- tensor --- */

enum StorageMode {
    columnMajor,
    rowMajor,
    defaultMajor = columnMajor
};

namespace gputils_detail {
template <typename T>
inline constexpr bool is_fp = std::is_same_v<T, float> || std::is_same_v<T, double>;

template <typename T>
inline void require_fp() {
    static_assert(is_fp<T>, "this DTensor operation needs T = float or double");
}
}  // namespace gputils_detail

/**
 * (m, n, k) tensor of k column-major m x n matrices on the GPU
 * (tensor.cuh:271-661).  Owns its data unless it is a slice or the target of
 * operator= (which aliases).  For k > 1 a device array of per-matrix
 * pointers is kept, as in the reference, so that ptrMatrices() and the byte
 * count match.
 */
TEMPLATE_WITH_TYPE_T
class DTensor {
private:
    T *m_d_data = nullptr;
    T **m_d_ptrMatrices = nullptr;
    size_t m_numRows = 0;
    size_t m_numCols = 0;
    size_t m_numMats = 0;
    bool m_doDestroyData = false;
    bool m_doDestroyPtrMatrices = false;
    size_t m_ptrCapacity = 0;  ///< entries of the owned pointer array
    size_t m_idxStream = 0;

    void destroy() {
        if (m_doDestroyData) {
            if (m_d_data) gpuErrChk(gputils_free(m_d_data));
            m_d_data = nullptr;
            Session::getInstance().incrementAllocatedBytes(-(long long)(m_numRows * m_numCols * m_numMats * sizeof(T)));
            m_doDestroyData = false;
        }
        if (m_doDestroyPtrMatrices) {
            if (m_d_ptrMatrices) gpuErrChk(gputils_free(m_d_ptrMatrices));
            m_d_ptrMatrices = nullptr;
            Session::getInstance().incrementAllocatedBytes(-(long long)(m_ptrCapacity * sizeof(T *)));
            m_doDestroyPtrMatrices = false;
            m_ptrCapacity = 0;
        }
    }

    void allocateOnDevice(size_t size, bool zero = false) {
        if (size == 0) return;
        destroy();
        m_doDestroyData = true;
        gpuErrChk(Session::getInstance().allocate((void **)&m_d_data, size * sizeof(T)));
        if (zero) gpuErrChk(gputils_memset(m_d_data, 0, size * sizeof(T)));
        if (m_numMats > 1) {
            m_doDestroyPtrMatrices = true;
            int st = Session::getInstance().allocate((void **)&m_d_ptrMatrices, m_numMats * sizeof(T *));
            if (st != GPUTILS_OK) {
                gputils_free(m_d_data);
                gpuErrChk(st);
            }
            m_ptrCapacity = m_numMats;
        } else {
            m_doDestroyPtrMatrices = false;
        }
    }

    void initialisePointersToMatricesData() {
        if (m_numMats <= 1 || !m_d_ptrMatrices || !m_doDestroyPtrMatrices) return;
        gpuErrChk(gputils_fill_pointer_array((void **)m_d_ptrMatrices, m_d_data,
                                             (int64_t)(m_numRows * m_numCols * sizeof(T)),
                                             (int64_t)m_numMats, stream()));
        gpuErrChk(gputils_stream_synchronize(stream()));
    }

    void rm2cm(const std::vector<T> &rm, std::vector<T> &cm) const {
        const size_t n = m_numRows * m_numCols;
        for (size_t k = 0; k < m_numMats; k++) {
            const size_t idx = k * n;
            for (size_t r = 0; r < m_numRows; r++)
                for (size_t c = 0; c < m_numCols; c++) cm[idx + (r + c * m_numRows)] = rm[idx + (c + r * m_numCols)];
        }
    }

    std::ostream &print(std::ostream &out) const {
        out << "Tensor [" << m_numRows << " x " << m_numCols << " x " << m_numMats << "]:" << std::endl;
        std::vector<T> temp;
        download(temp);
        for (size_t k = 0; k < m_numMats; k++) {
            out << ">> layer: " << k << std::endl;
            for (size_t i = 0; i < m_numRows; i++) {
                for (size_t j = 0; j < m_numCols; j++)
                    out << std::setw(10) << temp[m_numRows * (m_numCols * k + j) + i] << ", ";
                out << std::endl;
            }
        }
        return out;
    }

    template <typename S>
    S reduceToHost(int (*fn)(int64_t, const S *, S *, void *, void *)) const;

public:
    /** Stream handle used by operations on this tensor. */
    void *stream() const { return Session::getInstance().stream(m_idxStream); }

    /** Set the stream index; returns a copy by value like the reference (tensor.cuh:663-670). */
    DTensor<T> setStreamIdx(size_t idx) {
        if (idx >= s_numStreams) throw std::invalid_argument("Invalid stream index; it exceeds the max allocated streams");
        m_idxStream = idx;
        return *this;
    }

    size_t streamIdx() const { return m_idxStream; }

    static DTensor<T> createRandomTensor(size_t numRows, size_t numCols, size_t numMats, T low, T hi);
    static DTensor<T> parseFromFile(std::string path_to_file, StorageMode mode = StorageMode::defaultMajor);

    DTensor() = default;
    ~DTensor() { destroy(); }

    /** (m,n,k) tensor, optionally zeroed (tensor.cuh:863-872). */
    DTensor(size_t m, size_t n = 1, size_t k = 1, bool zero = false) {
        m_numRows = m;
        m_numCols = n;
        m_numMats = k;
        allocateOnDevice(m * n * k, zero);
        initialisePointersToMatricesData();
    }

    /** (m,n,k) tensor uploaded from host data (tensor.cuh:874-884). */
    DTensor(const std::vector<T> &data, size_t m, size_t n = 1, size_t k = 1,
            StorageMode mode = StorageMode::defaultMajor) {
        m_numRows = m;
        m_numCols = n;
        m_numMats = k;
        allocateOnDevice(m * n * k);
        upload(data, mode);
        initialisePointersToMatricesData();
    }

    /** Deep copy (tensor.cuh:886-897). */
    DTensor(const DTensor &other) {
        m_numMats = other.m_numMats;
        m_numRows = other.m_numRows;
        m_numCols = other.m_numCols;
        m_idxStream = other.m_idxStream;
        allocateOnDevice(m_numRows * m_numCols * m_numMats);
        Session::getInstance().synchronizeAllStreams();
        gpuErrChk(gputils_memcpy(m_d_data, other.raw(), numEl() * sizeof(T), 2));
        initialisePointersToMatricesData();
    }

    /** Move (tensor.cuh:927-946). */
    DTensor(DTensor &&other) {
        m_numCols = other.m_numCols;
        m_numRows = other.m_numRows;
        m_numMats = other.m_numMats;
        m_d_data = other.m_d_data;
        m_doDestroyData = other.m_doDestroyData;
        m_doDestroyPtrMatrices = other.m_doDestroyPtrMatrices;
        m_d_ptrMatrices = other.m_d_ptrMatrices;
        m_ptrCapacity = other.m_ptrCapacity;
        m_idxStream = other.m_idxStream;
        other.m_ptrCapacity = 0;
        other.m_doDestroyPtrMatrices = false;
        other.m_doDestroyData = false;
        other.m_d_ptrMatrices = nullptr;
        other.m_d_data = nullptr;
        other.m_numCols = 0;
        other.m_numRows = 0;
        other.m_numMats = 0;
    }

    /**
     * Slice (tensor.cuh:899-925), `to` inclusive.  axis 2: matrices
     * from..to (shares the parent's pointer array); axis 1: columns of the
     * first matrix; axis 0: flat element range as a column vector.
     */
    DTensor(const DTensor &other, size_t axis, size_t from, size_t to) {
        if (from > to) throw std::invalid_argument("from > to");
        size_t offset = 0, len = to - from + 1;
        m_d_ptrMatrices = nullptr;
        if (axis == 2) {
            offset = other.m_numRows * other.m_numCols * from;
            m_numRows = other.m_numRows;
            m_numCols = other.m_numCols;
            m_numMats = len;
            m_d_ptrMatrices = other.m_d_ptrMatrices ? other.m_d_ptrMatrices + from : nullptr;
        } else if (axis == 1) {
            offset = other.m_numRows * from;
            m_numRows = other.m_numRows;
            m_numCols = len;
            m_numMats = 1;
        } else if (axis == 0) {
            offset = from;
            m_numRows = len;
            m_numCols = 1;
            m_numMats = 1;
        }
        m_d_data = other.m_d_data + offset;
        m_doDestroyData = false;
        m_doDestroyPtrMatrices = false;
        m_idxStream = other.m_idxStream;
    }

    T *raw() const { return m_d_data; }
    T **ptrMatrices() const { return m_d_ptrMatrices; }
    size_t numRows() const { return m_numRows; }
    size_t numCols() const { return m_numCols; }
    size_t numMats() const { return m_numMats; }
    size_t numEl() const { return m_numRows * m_numCols * m_numMats; }

    /** Synchronous upload (tensor.cuh:1128-1145). */
    bool upload(const std::vector<T> &vec, StorageMode mode = StorageMode::defaultMajor) {
        const size_t size = vec.size();
        const size_t thisSize = numEl();
        if (size != thisSize) throw std::invalid_argument("[upload] vec has wrong size");
        Session::getInstance().synchronizeAllStreams();
        if (mode == StorageMode::rowMajor) {
            std::vector<T> vecCm(thisSize);
            rm2cm(vec, vecCm);
            gpuErrChk(gputils_memcpy(m_d_data, vecCm.data(), size * sizeof(T), 0));
        } else {
            gpuErrChk(gputils_memcpy(m_d_data, vec.data(), size * sizeof(T), 0));
        }
        return true;
    }

    /** Synchronous download of all elements (tensor.cuh:1147-1154). */
    void download(std::vector<T> &vec) const {
        vec.resize(numEl());
        Session::getInstance().synchronizeAllStreams();
        gpuErrChk(gputils_memcpy(vec.data(), m_d_data, numEl() * sizeof(T), 1));
    }

    /** Device-to-device copy into a tensor with at least numEl() elements (tensor.cuh:1199-1208). */
    void deviceCopyTo(DTensor<T> &elsewhere) const {
        if (elsewhere.numEl() < numEl()) throw std::invalid_argument("[deviceCopyTo] tensor does not fit into destination");
        Session::getInstance().synchronizeAllStreams();
        gpuErrChk(gputils_memcpy(elsewhere.raw(), m_d_data, numEl() * sizeof(T), 2));
    }

    /** Rows rowsFrom..rowsTo (inclusive) of matrix matIdx as a new tensor (tensor.cuh:1396-1424). */
    DTensor<T> getRows(size_t rowsFrom, size_t rowsTo, size_t matIdx) const {
        const size_t len = rowsTo - rowsFrom + 1;
        DTensor<T> rowsOnly(len, m_numCols, 1);
        const T *src = m_d_data + matIdx * m_numRows * m_numCols + rowsFrom;
        // rows of a column-major matrix: one contiguous copy per column
        Session::getInstance().synchronizeAllStreams();
        for (size_t c = 0; c < m_numCols; c++)
            gpuErrChk(gputils_memcpy(rowsOnly.raw() + c * len, src + c * m_numRows, len * sizeof(T), 2));
        return rowsOnly;
    }

    /** Batched transpose (tensor.cuh:1167-1197, one launch for all matrices). */
    DTensor<T> tr() const {
        gputils_detail::require_fp<T>();
        DTensor<T> transposes(m_numCols, m_numRows, m_numMats);
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dtranspose_batched(m_numRows, m_numCols, raw(), m_numRows, m_numRows * m_numCols,
                                                 transposes.raw(), m_numCols, m_numRows * m_numCols, m_numMats, stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_stranspose_batched(m_numRows, m_numCols, raw(), m_numRows, m_numRows * m_numCols,
                                                 transposes.raw(), m_numCols, m_numRows * m_numCols, m_numMats, stream()));
        return transposes;
    }

    /** Frobenius dot product (tensor.cuh:968-992). */
    T dotF(const DTensor &other);
    /** Frobenius norm (tensor.cuh:994-1010). */
    T normF() const;
    /** Sum of |x_i| (tensor.cuh:1012-1028). */
    T sumAbs() const;
    /** max |x_i| (tensor.cuh:1030-1050). */
    T maxAbs() const;
    /** min |x_i| (tensor.cuh:1052-1072). */
    T minAbs() const;

    /**
     * Batched least squares b_i <- argmin ||A_i x - b_i|| for m >= n, nrhs 1;
     * the solution overwrites the first n rows of b_i and A_i is
     * overwritten with its QR factors (tensor.cuh:1340-1394).
     */
    void leastSquaresBatched(DTensor &b);

    /** C_i <- beta C_i + alpha A_i B_i (tensor.cuh:1286-1338). */
    void addAB(const DTensor<T> &A, const DTensor<T> &B, T alpha = 1, T beta = 0);

    /**
     * Reshape keeping the element count (tensor.cuh:828-861).  The pointer
     * array is reallocated when the matrix count outgrows it; unlike the
     * reference, the byte counter always subtracts exactly what was
     * allocated.
     */
    void reshape(size_t newNumRows, size_t newNumCols, size_t newNumMats = 1) {
        if (m_numRows == newNumRows && m_numCols == newNumCols && m_numMats == newNumMats) return;
        const size_t newNumElements = newNumRows * newNumCols * newNumMats;
        if (numEl() != newNumElements) {
            char errMessage[256];
            snprintf(errMessage, sizeof(errMessage),
                     "DTensor[%zu x %zu x %zu] with %zu elements cannot be reshaped into DTensor[%zu x %zu x %zu] (%zu elements)",
                     numRows(), numCols(), numMats(), numEl(), newNumRows, newNumCols, newNumMats, newNumElements);
            throw std::invalid_argument(errMessage);
        }
        if (newNumMats > m_numMats && newNumMats > m_ptrCapacity) {
            if (m_d_ptrMatrices && m_doDestroyPtrMatrices) {
                gpuErrChk(gputils_free(m_d_ptrMatrices));
                Session::getInstance().incrementAllocatedBytes(-(long long)(m_ptrCapacity * sizeof(T *)));
                m_d_ptrMatrices = nullptr;
                m_doDestroyPtrMatrices = false;
                m_ptrCapacity = 0;
            }
            if (newNumMats > 1) {
                gpuErrChk(Session::getInstance().allocate((void **)&m_d_ptrMatrices, newNumMats * sizeof(T *)));
                m_doDestroyPtrMatrices = true;
                m_ptrCapacity = newNumMats;
            }
        }
        m_numRows = newNumRows;
        m_numCols = newNumCols;
        m_numMats = newNumMats;
        initialisePointersToMatricesData();
    }

    /** Save as text, or binary .bt (tensor.cuh:797-825). */
    void saveToFile(std::string pathToFile);

    /**
     * Columns i and j rotated: col_i <- c col_i + s col_j,
     * col_j <- c col_j - s col_i with s = *minus_s (tensor.cuh:1074-1086;
     * c and minus_s are host pointers).  Matrices only (numMats = 1);
     * i, j < numCols.
     */
    void applyRightGivensRotation(size_t i, size_t j, const T *c, const T *minus_s);

    /** Rows i and j rotated likewise (tensor.cuh:1088-1104); i, j < numRows. */
    void applyLeftGivensRotation(size_t i, size_t j, const T *c, const T *minus_s);

    /* ------------- operators ------------- */

    /** Shallow alias (tensor.cuh:1219-1228). */
    DTensor &operator=(const DTensor &other) {
        m_numMats = other.m_numMats;
        m_numRows = other.m_numRows;
        m_numCols = other.m_numCols;
        m_doDestroyData = false;
        m_d_data = other.m_d_data;
        m_idxStream = other.m_idxStream;
        return *this;
    }

    /** Element (i,j,k) by a synchronous single-element copy (tensor.cuh:1278-1284). */
    T operator()(size_t i, size_t j = 0, size_t k = 0) const {
        T hostDst;
        const size_t offset = i + m_numRows * (j + m_numCols * k);
        Session::getInstance().synchronizeAllStreams();
        gpuErrChk(gputils_memcpy(&hostDst, m_d_data + offset, sizeof(T), 1));
        return hostDst;
    }

    DTensor &operator*=(T scalar) {
        gputils_detail::require_fp<T>();
        if constexpr (std::is_same_v<T, double>) gpuErrChk(gputils_dscal(numEl(), scalar, m_d_data, stream()));
        if constexpr (std::is_same_v<T, float>) gpuErrChk(gputils_sscal(numEl(), scalar, m_d_data, stream()));
        return *this;
    }

    DTensor &operator+=(const DTensor &rhs) {
        gputils_detail::require_fp<T>();
        if constexpr (std::is_same_v<T, double>) gpuErrChk(gputils_daxpy(numEl(), 1.0, rhs.m_d_data, m_d_data, stream()));
        if constexpr (std::is_same_v<T, float>) gpuErrChk(gputils_saxpy(numEl(), 1.0f, rhs.m_d_data, m_d_data, stream()));
        return *this;
    }

    DTensor &operator-=(const DTensor &rhs) {
        gputils_detail::require_fp<T>();
        if constexpr (std::is_same_v<T, double>) gpuErrChk(gputils_daxpy(numEl(), -1.0, rhs.m_d_data, m_d_data, stream()));
        if constexpr (std::is_same_v<T, float>) gpuErrChk(gputils_saxpy(numEl(), -1.0f, rhs.m_d_data, m_d_data, stream()));
        return *this;
    }

    friend DTensor operator+(DTensor &first, const DTensor &second) {
        DTensor result(first);
        result += second;
        return result;
    }

    friend DTensor operator-(DTensor &first, const DTensor &second) {
        DTensor result(first);
        result -= second;
        return result;
    }

    friend DTensor<T> operator*(DTensor &A, DTensor &B) {
        DTensor<T> result(A.m_numRows, B.m_numCols, B.m_numMats);
        result.addAB(A, B);
        return result;
    }

    friend DTensor<T> operator*(T a, DTensor &B) {
        DTensor<T> result(B);
        result *= a;
        return result;
    }

    friend std::ostream &operator<<(std::ostream &out, const DTensor<T> &data) { return data.print(out); }
};

/* --- reductions: one device result copied back on the tensor's stream --- */

template <typename T>
template <typename S>
S DTensor<T>::reduceToHost(int (*fn)(int64_t, const S *, S *, void *, void *)) const {
    Session &s = Session::getInstance();
    S *slot = static_cast<S *>(s.scalarSlot(m_idxStream));
    gpuErrChk(fn((int64_t)numEl(), m_d_data, slot, s.reduceWorkspace(m_idxStream), stream()));
    S host;
    gpuErrChk(gputils_memcpy_async(&host, slot, sizeof(S), 1, stream()));
    gpuErrChk(gputils_stream_synchronize(stream()));
    return host;
}

namespace gputils_detail {
template <typename T>
struct Reducers;
template <>
struct Reducers<double> {
    static int nrm2(int64_t n, const double *x, double *r, void *w, void *s) { return gputils_dnrm2(n, x, r, w, s); }
    static int asum(int64_t n, const double *x, double *r, void *w, void *s) { return gputils_dasum(n, x, r, w, s); }
    static int amax(int64_t n, const double *x, double *r, void *w, void *s) { return gputils_damax(n, x, r, w, s); }
    static int amin(int64_t n, const double *x, double *r, void *w, void *s) { return gputils_damin(n, x, r, w, s); }
};
template <>
struct Reducers<float> {
    static int nrm2(int64_t n, const float *x, float *r, void *w, void *s) { return gputils_snrm2(n, x, r, w, s); }
    static int asum(int64_t n, const float *x, float *r, void *w, void *s) { return gputils_sasum(n, x, r, w, s); }
    static int amax(int64_t n, const float *x, float *r, void *w, void *s) { return gputils_samax(n, x, r, w, s); }
    static int amin(int64_t n, const float *x, float *r, void *w, void *s) { return gputils_samin(n, x, r, w, s); }
};
}  // namespace gputils_detail

template <typename T>
T DTensor<T>::dotF(const DTensor &other) {
    gputils_detail::require_fp<T>();
    if (m_numRows != other.m_numRows || m_numCols != other.m_numCols || m_numMats != other.m_numMats)
        throw std::invalid_argument("[dotF] incompatible dimensions");
    Session &s = Session::getInstance();
    T *slot = static_cast<T *>(s.scalarSlot(m_idxStream));
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_ddot(numEl(), m_d_data, other.m_d_data, slot, s.reduceWorkspace(m_idxStream), stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_sdot(numEl(), m_d_data, other.m_d_data, slot, s.reduceWorkspace(m_idxStream), stream()));
    T host;
    gpuErrChk(gputils_memcpy_async(&host, slot, sizeof(T), 1, stream()));
    gpuErrChk(gputils_stream_synchronize(stream()));
    return host;
}

template <typename T>
T DTensor<T>::normF() const {
    gputils_detail::require_fp<T>();
    return reduceToHost<T>(&gputils_detail::Reducers<T>::nrm2);
}
template <typename T>
T DTensor<T>::sumAbs() const {
    gputils_detail::require_fp<T>();
    return reduceToHost<T>(&gputils_detail::Reducers<T>::asum);
}
template <typename T>
T DTensor<T>::maxAbs() const {
    gputils_detail::require_fp<T>();
    return reduceToHost<T>(&gputils_detail::Reducers<T>::amax);
}
template <typename T>
T DTensor<T>::minAbs() const {
    gputils_detail::require_fp<T>();
    return reduceToHost<T>(&gputils_detail::Reducers<T>::amin);
}

template <typename T>
void DTensor<T>::addAB(const DTensor<T> &A, const DTensor<T> &B, T alpha, T beta) {
    gputils_detail::require_fp<T>();
    const size_t nMat = A.numMats(), nRA = A.numRows(), nCA = A.numCols(), nCB = B.numCols();
    /* like the reference: sizes from A and B, k from A, no further checks */
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_dgemm_strided_batched(nRA, nCB, nCA, alpha, A.raw(), nRA, nRA * nCA, B.raw(), nCA, nCA * nCB,
                                                beta, raw(), nRA, nRA * nCB, nMat, stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_sgemm_strided_batched(nRA, nCB, nCA, alpha, A.raw(), nRA, nRA * nCA, B.raw(), nCA, nCA * nCB,
                                                beta, raw(), nRA, nRA * nCB, nMat, GPUTILS_GEMM_FP32, stream()));
}

/**
 * Batched least squares (tensor.cuh:1340-1394): for every i, b_i[0:n] <-
 * argmin ||A_i x - b_i|| (m >= n, one right-hand side).  A_i is overwritten
 * with its Householder QR factors.
 */
template <typename T>
void DTensor<T>::leastSquaresBatched(DTensor &B) {
    gputils_detail::require_fp<T>();
    const size_t batchSize = numMats(), nColsB = B.numCols();
    if (B.numRows() != m_numRows)
        throw std::invalid_argument("[Least squares batched] rhs rows does not equal lhs rows");
    if (nColsB != 1) throw std::invalid_argument("[Least squares batched] rhs are not vectors");
    if (B.numMats() != batchSize)
        throw std::invalid_argument("[Least squares batched] rhs numMats does not equal lhs numMats");
    if (m_numCols > m_numRows)
        throw std::invalid_argument("[Least squares batched] supports square or tall matrices only");
    const int64_t m = (int64_t)m_numRows, n = (int64_t)m_numCols;
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_dgels_batched(m, n, 1, raw(), m, m * n, B.raw(), m, m, nullptr, (int64_t)batchSize, stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_sgels_batched(m, n, 1, raw(), m, m * n, B.raw(), m, m, nullptr, (int64_t)batchSize, stream()));
}

template <typename T>
DTensor<T> DTensor<T>::createRandomTensor(size_t numRows, size_t numCols, size_t numMats, T low, T hi) {
    if constexpr (std::is_floating_point<T>::value) {
        auto randVec = generateRealRandomVector<T>(numRows * numCols * numMats, low, hi);
        return DTensor<T>(randVec, numRows, numCols, numMats);
    } else if constexpr (std::is_same_v<T, int>) {
        auto randVec = generateIntRandomVector(numRows * numCols * numMats, low, hi);
        return DTensor<T>(randVec, numRows, numCols, numMats);
    }
    throw std::invalid_argument("[createRandomTensor] unsupported type T");
}

/* --------------------------------------------------------- file I/O --- */

template <typename T>
struct data_t {
    size_t numRows;
    size_t numCols;
    size_t numMats;
    std::vector<T> data;
};

/** Text format: rows, cols, mats, then one value per line (tensor.cuh:713-760). */
template <typename T>
data_t<T> vectorFromTextFile(std::string path_to_file) {
    data_t<T> d;
    std::ifstream file(path_to_file, std::ios::in);
    if (!file.is_open()) throw std::invalid_argument("[vectorFromTextFile] the file does not exist");
    std::string line;
    getline(file, line);
    d.numRows = atoi(line.c_str());
    getline(file, line);
    d.numCols = atoi(line.c_str());
    getline(file, line);
    d.numMats = atoi(line.c_str());
    const size_t numElements = d.numRows * d.numCols * d.numMats;
    std::vector<T> v(numElements);
    size_t i = 0;
    while (i < numElements && getline(file, line)) {
        if constexpr (std::is_same_v<T, int>) v[i] = atoi(line.c_str());
        else if constexpr (std::is_same_v<T, double>) v[i] = std::stod(line);
        else if constexpr (std::is_same_v<T, float>) v[i] = std::stof(line);
        else if constexpr (std::is_same_v<T, long double>) v[i] = std::stold(line);
        else if constexpr (std::is_same_v<T, long>) v[i] = std::stol(line);
        else if constexpr (std::is_same_v<T, long long>) v[i] = std::stoll(line);
        else if constexpr (std::is_same_v<T, unsigned long>) v[i] = std::stoul(line);
        else if constexpr (std::is_same_v<T, unsigned long long>) v[i] = std::stoull(line);
        else throw std::invalid_argument("data type not supported");
        ++i;
    }
    d.data = v;
    return d;
}

/** Binary .bt: three little-endian uint64 (rows, cols, mats) then the data (tensor.cuh:762-782). */
template <typename T>
data_t<T> vectorFromBinaryFile(std::string path_to_file) {
    data_t<T> d;
    std::ifstream in(path_to_file, std::ios::binary);
    if (!in.is_open()) throw std::invalid_argument("[vectorFromBinaryFile] the file does not exist");
    uint64_t hdr[3] = {0, 0, 0};
    in.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    d.numRows = hdr[0];
    d.numCols = hdr[1];
    d.numMats = hdr[2];
    const uint64_t numElements = hdr[0] * hdr[1] * hdr[2];
    d.data.resize(numElements);
    in.read(reinterpret_cast<char *>(d.data.data()), numElements * sizeof(T));
    return d;
}

template <typename T>
DTensor<T> DTensor<T>::parseFromFile(std::string path_to_file, StorageMode mode) {
    const bool binary = path_to_file.size() >= 3 && path_to_file.substr(path_to_file.size() - 3) == ".bt";
    data_t<T> parsed = binary ? vectorFromBinaryFile<T>(path_to_file) : vectorFromTextFile<T>(path_to_file);
    return DTensor<T>(parsed.data, parsed.numRows, parsed.numCols, parsed.numMats, mode);
}

template <typename T>
void DTensor<T>::applyRightGivensRotation(size_t i, size_t j, const T *c, const T *minus_s) {
    gputils_detail::require_fp<T>();
    if (m_numMats > 1) throw std::invalid_argument("[applyRightGivensRotation] tensors (nMat>1) not supported");
    if (i >= m_numCols || j >= m_numCols) throw std::invalid_argument("[applyRightGivensRotation] invalid column index");
    T *col_i = m_d_data + i * m_numRows;
    T *col_j = m_d_data + j * m_numRows;
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_drot((int64_t)m_numRows, col_i, 1, col_j, 1, *c, *minus_s, stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_srot((int64_t)m_numRows, col_i, 1, col_j, 1, *c, *minus_s, stream()));
}

template <typename T>
void DTensor<T>::applyLeftGivensRotation(size_t i, size_t j, const T *c, const T *minus_s) {
    gputils_detail::require_fp<T>();
    if (m_numMats > 1) throw std::invalid_argument("[applyLeftGivensRotation] tensors (nMat>1) not supported");
    if (i >= m_numRows || j >= m_numRows) throw std::invalid_argument("[applyLeftGivensRotation] invalid row index");
    const int64_t ld = (int64_t)m_numRows;
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_drot((int64_t)m_numCols, m_d_data + i, ld, m_d_data + j, ld, *c, *minus_s, stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_srot((int64_t)m_numCols, m_d_data + i, ld, m_d_data + j, ld, *c, *minus_s, stream()));
}

template <typename T>
void DTensor<T>::saveToFile(std::string pathToFile) {
    std::vector<T> myData(numEl());
    download(myData);
    const bool binary = pathToFile.size() >= 3 && pathToFile.substr(pathToFile.size() - 3) == ".bt";
    if (binary) {
        const uint64_t hdr[3] = {(uint64_t)numRows(), (uint64_t)numCols(), (uint64_t)numMats()};
        std::ofstream out(pathToFile, std::ios::binary);
        out.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
        out.write(reinterpret_cast<const char *>(myData.data()), myData.size() * sizeof(T));
    } else {
        std::ofstream file(pathToFile);
        file << numRows() << std::endl << numCols() << std::endl << numMats() << std::endl;
        if constexpr (std::is_floating_point<T>::value) file << std::setprecision(std::numeric_limits<T>::max_digits10);
        for (const T &el : myData) file << el << std::endl;
    }
}

/* ------------------------------------------------------------ status --- */

/** Device info codes, one per matrix (tensor.cuh:1453-1469). */
class IStatus {
protected:
    std::unique_ptr<DTensor<int>> m_info;
    explicit IStatus(size_t n = 1) { m_info = std::make_unique<DTensor<int>>(1, 1, n, true); }

public:
    virtual ~IStatus() = default;
    virtual DTensor<int> &info() { return *m_info; }
};

/* ---------------------------------------------------------- Cholesky --- */

/**
 * Batched lower Cholesky factorisation and solve (tensor.cuh:2098-2197).
 * A is either the matrices to factorise or (factorised = true) their lower
 * factors.  A must outlive this object.
 */
TEMPLATE_WITH_TYPE_T
class CholeskyBatchFactoriser : public IStatus {
private:
    DTensor<T> *m_matrix;
    size_t m_numRows = 0;
    size_t m_numMats = 0;
    bool m_factorisationDone = false;

public:
    CholeskyBatchFactoriser() = delete;

    CholeskyBatchFactoriser(DTensor<T> &A, bool factorised = false) : IStatus(A.numMats()), m_factorisationDone(factorised) {
        gputils_detail::require_fp<T>();
        if (A.numRows() != A.numCols()) throw std::invalid_argument("[CholeskyBatch] A must be square");
        m_matrix = &A;
        m_numRows = A.numRows();
        m_numMats = A.numMats();
    }

    /** Factorise all matrices in place (idempotent, tensor.cuh:2135-2159). */
    void factorise() {
        if (m_factorisationDone) return;
        const int64_t n = (int64_t)m_numRows;
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dpotrf_batched(n, m_matrix->raw(), n, n * n, m_info->raw(), m_numMats, m_matrix->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_spotrf_batched(n, m_matrix->raw(), n, n * n, m_info->raw(), m_numMats, m_matrix->stream()));
        m_factorisationDone = true;
    }

    /**
     * Solve A_i x_i = b_i in place (tensor.cuh:2161-2197).  The reference
     * requires b with one column; any number of columns is accepted here.
     */
    void solve(DTensor<T> &b) {
        if (!m_factorisationDone) throw std::logic_error("[CholeskyBatchSolve] no factor to solve with");
        if (m_numRows != b.numRows() || m_numMats != b.numMats())
            throw std::invalid_argument("[CholeskyBatchSolve] A and b incompatible");
        if (b.numCols() < 1) throw std::invalid_argument("[CholeskyBatchSolve] b has no columns");
        const int64_t n = (int64_t)m_numRows, nrhs = (int64_t)b.numCols();
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dpotrs_batched(n, nrhs, m_matrix->raw(), n, n * n, b.raw(), n, n * nrhs, m_info->raw(),
                                             m_numMats, m_matrix->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_spotrs_batched(n, nrhs, m_matrix->raw(), n, n * n, b.raw(), n, n * nrhs, m_info->raw(),
                                             m_numMats, m_matrix->stream()));
    }
};

/**
 * Single-matrix Cholesky (tensor.cuh:1689-1783): the k = 1 case of the
 * batched kernels.
 */
TEMPLATE_WITH_TYPE_T
class CholeskyFactoriser : public IStatus {
private:
    DTensor<T> *m_matrix;

public:
    CholeskyFactoriser(DTensor<T> &A) : IStatus() {
        gputils_detail::require_fp<T>();
        if (A.numMats() > 1) throw std::invalid_argument("[Cholesky] 3D tensors require `CholeskyBatchFactoriser`");
        if (A.numRows() != A.numCols()) throw std::invalid_argument("[Cholesky] Matrix A must be square");
        m_matrix = &A;
    }

    void factorise() {
        const int64_t n = (int64_t)m_matrix->numRows();
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dpotrf_batched(n, m_matrix->raw(), n, n * n, m_info->raw(), 1, m_matrix->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_spotrf_batched(n, m_matrix->raw(), n, n * n, m_info->raw(), 1, m_matrix->stream()));
    }

    /** Solves A x = b for the single column b (the reference passes nrhs = 1). */
    void solve(DTensor<T> &b) {
        const int64_t n = (int64_t)m_matrix->numRows();
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dpotrs_batched(n, 1, m_matrix->raw(), n, n * n, b.raw(), n, n, m_info->raw(), 1,
                                             m_matrix->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_spotrs_batched(n, 1, m_matrix->raw(), n, n * n, b.raw(), n, n, m_info->raw(), 1,
                                             m_matrix->stream()));
    }
};

/* ---------------------------------------------------------------- QR --- */

/**
 * Householder QR of one tall or square matrix (tensor.cuh:1796-1995):
 * factorise() overwrites A with R and the reflectors (LAPACK storage),
 * leastSquares(b) overwrites b[0:n] with argmin ||A x - b||, getQR fills
 * the thin Q (m x n) and R (n x n).  A must outlive this object.
 */
TEMPLATE_WITH_TYPE_T
class QRFactoriser : public IStatus {
private:
    std::unique_ptr<DTensor<T>> m_householder;  ///< tau, n values
    DTensor<T> *m_matrix;

public:
    QRFactoriser(DTensor<T> &A) : IStatus() {
        gputils_detail::require_fp<T>();
        if (A.numMats() > 1) throw std::invalid_argument("[QR] 3D tensors require `leastSquaresBatched`");
        if (A.numRows() < A.numCols()) throw std::invalid_argument("[QR] Matrix A must be tall or square");
        m_matrix = &A;
        m_householder = std::make_unique<DTensor<T>>(A.numCols());
    }

    void factorise() {
        const int64_t m = (int64_t)m_matrix->numRows(), n = (int64_t)m_matrix->numCols();
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dgeqrf_batched(m, n, m_matrix->raw(), m, m * n, m_householder->raw(), n, m_info->raw(), 1,
                                             m_matrix->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_sgeqrf_batched(m, n, m_matrix->raw(), m, m * n, m_householder->raw(), n, m_info->raw(), 1,
                                             m_matrix->stream()));
    }

    /** b <- Q^T b, then b[0:n] <- R^{-1} b[0:n] (tensor.cuh:1891-1927). */
    void leastSquares(DTensor<T> &b) {
        const int64_t m = (int64_t)m_matrix->numRows(), n = (int64_t)m_matrix->numCols();
        if (b.numRows() != (size_t)m) throw std::invalid_argument("[QR] b must have as many rows as A");
        if constexpr (std::is_same_v<T, double>) {
            gpuErrChk(gputils_dormqr_batched(1, m, 1, n, m_matrix->raw(), m, m * n, m_householder->raw(), n, b.raw(), m,
                                             m, 1, m_matrix->stream()));
            gpuErrChk(gputils_dtrsm_upper_batched(n, 1, m_matrix->raw(), m, m * n, b.raw(), m, m, 1, m_matrix->stream()));
        }
        if constexpr (std::is_same_v<T, float>) {
            gpuErrChk(gputils_sormqr_batched(1, m, 1, n, m_matrix->raw(), m, m * n, m_householder->raw(), n, b.raw(), m,
                                             m, 1, m_matrix->stream()));
            gpuErrChk(gputils_strsm_upper_batched(n, 1, m_matrix->raw(), m, m * n, b.raw(), m, m, 1, m_matrix->stream()));
        }
    }

    /** Thin Q = H_0 ... H_{n-1} [I; 0] and R = triu(A[0:n, :]) (tensor.cuh:1929-1995). */
    void getQR(DTensor<T> &Q, DTensor<T> &R) {
        const size_t m = m_matrix->numRows(), n = m_matrix->numCols();
        if (Q.numRows() != m || Q.numCols() != n) throw std::invalid_argument("[QR] invalid shape of Q.");
        if (R.numRows() != n || R.numCols() != n) throw std::invalid_argument("[QR] invalid shape of R.");
        std::vector<T> vecQ(m * n, T(0));
        for (size_t c = 0; c < n; c++) vecQ[c + c * m] = T(1);
        Q.upload(vecQ);
        const int64_t mm = (int64_t)m, nn = (int64_t)n;
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dormqr_batched(0, mm, nn, nn, m_matrix->raw(), mm, mm * nn, m_householder->raw(), nn, Q.raw(),
                                             mm, mm * nn, 1, m_matrix->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_sormqr_batched(0, mm, nn, nn, m_matrix->raw(), mm, mm * nn, m_householder->raw(), nn, Q.raw(),
                                             mm, mm * nn, 1, m_matrix->stream()));
        std::vector<T> qr(m * n);
        m_matrix->download(qr);
        std::vector<T> vecR(n * n, T(0));
        for (size_t c = 0; c < n; c++)
            for (size_t r = 0; r <= c; r++) vecR[r + c * n] = qr[r + c * m];
        R.upload(vecR);
    }
};

/* --------------------------------------------------------------- SVD --- */

/**
 * Batched SVD of tall matrices, A_i = U_i S_i V_i' (tensor.cuh:1500-1676).
 * All matrices are factorised by one launch (the reference loops over
 * them on the host).  At most 64 columns.  rank() counts sigma > epsilon
 * afresh on every call (the reference adds to a counter that is never
 * reset).  A is not modified by factorise(); with destroyMatrix = false a
 * private copy is factorised, as in the reference.
 */
TEMPLATE_WITH_TYPE_T
class Svd : public IStatus {
private:
    DTensor<T> *m_tensor = nullptr;
    std::unique_ptr<DTensor<T>> m_copy;
    std::shared_ptr<DTensor<T>> m_Vtr;
    std::shared_ptr<DTensor<T>> m_S;
    std::shared_ptr<DTensor<T>> m_U;
    std::shared_ptr<DTensor<unsigned int>> m_rank;
    void *m_work = nullptr;
    size_t m_workBytes = 0;
    bool m_computeU = false;

    void checkMatrix(DTensor<T> &tensor) const {
        if (tensor.numRows() < tensor.numCols()) throw std::invalid_argument("[svd] your matrix is fat (no offence)");
        if (tensor.numCols() > 64) throw std::invalid_argument("[svd] at most 64 columns are supported");
    }

public:
    Svd(DTensor<T> &mat, bool computeU = false, bool destroyMatrix = true) : IStatus(mat.numMats()) {
        gputils_detail::require_fp<T>();
        checkMatrix(mat);
        if (destroyMatrix) {
            m_tensor = &mat;
        } else {
            m_copy = std::make_unique<DTensor<T>>(mat);
            m_tensor = m_copy.get();
        }
        m_computeU = computeU;
        const size_t m = mat.numRows(), n = mat.numCols(), k = mat.numMats();
        m_Vtr = std::make_shared<DTensor<T>>(n, n, k);
        m_S = std::make_shared<DTensor<T>>(std::min(m, n), 1, k);
        m_rank = std::make_shared<DTensor<unsigned int>>(1, 1, k, true);
        if (computeU) m_U = std::make_shared<DTensor<T>>(m, m, k);
        const int64_t wb = gputils_gesvd_workspace_bytes((int)m, (int)n, (int64_t)k);
        if (wb < 0) throw std::invalid_argument("[svd] unsupported matrix size");
        if (wb > 0) {
            m_workBytes = (size_t)wb;
            gpuErrChk(Session::getInstance().allocate(&m_work, m_workBytes));
        }
    }

    ~Svd() {
        if (m_work) {
            gputils_free(m_work);
            Session::getInstance().incrementAllocatedBytes(-(long long)m_workBytes);
        }
    }

    Svd(const Svd &) = delete;
    Svd &operator=(const Svd &) = delete;

    void factorise() {
        const int m = (int)m_tensor->numRows(), n = (int)m_tensor->numCols();
        const int64_t k = (int64_t)m_tensor->numMats(), mn = (int64_t)m * n;
        T *U = m_computeU ? m_U->raw() : nullptr;
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dgesvd_batched(m, n, m_tensor->raw(), m, mn, m_S->raw(), n, U, m, (int64_t)m * m, m_Vtr->raw(),
                                             n, (int64_t)n * n, m_work, m_info->raw(), nullptr, k, m_tensor->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_sgesvd_batched(m, n, m_tensor->raw(), m, mn, m_S->raw(), n, U, m, (int64_t)m * m, m_Vtr->raw(),
                                             n, (int64_t)n * n, m_work, m_info->raw(), nullptr, k, m_tensor->stream()));
    }

    DTensor<T> &singularValues() const { return *m_S; }

    DTensor<T> const &rightSingularVectors() const { return *m_Vtr; }

    std::optional<std::shared_ptr<DTensor<T>>> leftSingularVectors() const {
        if (!m_computeU) return std::nullopt;
        return m_U;
    }

    DTensor<unsigned int> const &rank(T epsilon = 1e-6) const {
        const int n = (int)(m_S->numRows() * m_S->numCols());
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dsvd_rank_batched(n, m_S->raw(), n, epsilon, m_rank->raw(), (int64_t)m_rank->numMats(),
                                                m_tensor->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_ssvd_rank_batched(n, m_S->raw(), n, epsilon, m_rank->raw(), (int64_t)m_rank->numMats(),
                                                m_tensor->stream()));
        return *m_rank;
    }
};

/* --------------------------------------------------------- Nullspace --- */

/**
 * Nullspace of square or fat matrices (tensor.cuh:2008-2085): N_i spans
 * ker A_i (zero-padded to n x n), P_i = N_i N_i'.  Built by one SVD launch
 * of A', one rank count and one construction launch (the reference loops
 * over the matrices on the host).
 */
TEMPLATE_WITH_TYPE_T
class Nullspace {
private:
    std::unique_ptr<DTensor<T>> m_nullspace;
    std::unique_ptr<DTensor<T>> m_projOp;

public:
    Nullspace(DTensor<T> &a) {
        gputils_detail::require_fp<T>();
        const size_t m = a.numRows(), n = a.numCols(), nMats = a.numMats();
        if (m > n) throw std::invalid_argument("[nullspace] I was expecting a square or fat matrix");
        m_nullspace = std::make_unique<DTensor<T>>(n, n, nMats, true);
        m_projOp = std::make_unique<DTensor<T>>(n, n, nMats, true);
        if (m > 64) throw std::invalid_argument("[nullspace] at most 64 rows are supported");
        DTensor<T> aTranspose = a.tr();
        // SVD of a' (n x m) without right singular vectors: the sweeps run on
        // R^T and U comes from the accumulated rotations (same subspaces)
        DTensor<T> S(m, 1, nMats), U(n, n, nMats);
        DTensor<int> info(1, 1, nMats);
        DTensor<unsigned int> rank(1, 1, nMats, true);
        const int64_t nn = (int64_t)n * n, k = (int64_t)nMats;
        const int64_t wb = gputils_gesvd_workspace_bytes((int)n, (int)m, k);
        if (wb < 0) throw std::invalid_argument("[nullspace] unsupported matrix size");
        void *work = nullptr;
        if (wb > 0) gpuErrChk(Session::getInstance().allocate(&work, (size_t)wb));
        void *st = aTranspose.stream();
        if constexpr (std::is_same_v<T, double>) {
            gpuErrChk(gputils_dgesvd_batched((int)n, (int)m, aTranspose.raw(), n, (int64_t)n * m, S.raw(), m, U.raw(),
                                             n, nn, nullptr, m, (int64_t)m * m, work, info.raw(), nullptr, k, st));
            gpuErrChk(gputils_dsvd_rank_batched((int)m, S.raw(), m, 1e-6, rank.raw(), k, st));
            gpuErrChk(gputils_dnullspace_build_batched((int)n, U.raw(), n, nn, rank.raw(), m_nullspace->raw(), n, nn,
                                                       m_projOp->raw(), n, nn, k, st));
        }
        if constexpr (std::is_same_v<T, float>) {
            gpuErrChk(gputils_sgesvd_batched((int)n, (int)m, aTranspose.raw(), n, (int64_t)n * m, S.raw(), m, U.raw(),
                                             n, nn, nullptr, m, (int64_t)m * m, work, info.raw(), nullptr, k, st));
            gpuErrChk(gputils_ssvd_rank_batched((int)m, S.raw(), m, 1e-6f, rank.raw(), k, st));
            gpuErrChk(gputils_snullspace_build_batched((int)n, U.raw(), n, nn, rank.raw(), m_nullspace->raw(), n, nn,
                                                       m_projOp->raw(), n, nn, k, st));
        }
        gpuErrChk(gputils_stream_synchronize(st));
        if (work) {
            gputils_free(work);
            Session::getInstance().incrementAllocatedBytes(-wb);
        }
    }

    DTensor<T> const &nullspace() const { return *m_nullspace; }

    /** b_i <- P_i b_i in place (tensor.cuh:2081-2085). */
    void project(DTensor<T> &b) { b.addAB(*m_projOp, b, 1, 0); }
};

/* ------------------------------------------------------------ Givens --- */

/**
 * Zeroes element (k, j) of a matrix by a left Givens rotation of rows i and
 * k (tensor.cuh:2204-2312).  annihilate(i, k, j) is one kernel launch that
 * forms c = A(i,j) / hypot and s = A(k,j) / hypot on the device and
 * rotates the two rows (the reference launches a one-thread kernel for the
 * scalars, then cublas?rot in device pointer mode).  Index checks: i, k
 * < numRows and j < numCols (the reference checks i and k only).
 */
TEMPLATE_WITH_TYPE_T
class GivensAnnihilator {
private:
    DTensor<T> *m_matrix = nullptr;

public:
    GivensAnnihilator() { gputils_detail::require_fp<T>(); }

    GivensAnnihilator(DTensor<T> &a) {
        gputils_detail::require_fp<T>();
        setMatrix(a);
    }

    void setMatrix(DTensor<T> &a) {
        if (a.numMats() > 1) throw std::invalid_argument("[GivensAnnihilator] tensors (numMats > 1) not supported");
        m_matrix = &a;
    }

    void annihilate(size_t i, size_t k, size_t j) {
        if (!m_matrix) throw std::logic_error("[GivensAnnihilator::annihilate] no matrix set");
        const size_t nR = m_matrix->numRows(), nC = m_matrix->numCols();
        if (i >= nR || k >= nR) throw std::invalid_argument("[GivensAnnihilator::annihilate] invalid row index");
        if (j >= nC) throw std::invalid_argument("[GivensAnnihilator::annihilate] invalid column index j");
        const int64_t m = (int64_t)nR;
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dgivens_annihilate(m, (int64_t)nC, m_matrix->raw(), m, (int64_t)i, (int64_t)k,
                                                 (int64_t)j, m_matrix->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_sgivens_annihilate(m, (int64_t)nC, m_matrix->raw(), m, (int64_t)i, (int64_t)k,
                                                 (int64_t)j, m_matrix->stream()));
    }
};

/* ------------------------------------------------------------ graphs --- */

/**
 * Records the work that a callable enqueues on a Session stream into a HIP
 * graph and replays it (no reference counterpart: used to launch repeated
 * Nullspace::project calls without per-call launch overhead).  The
 * callable must only enqueue work on that stream (no host synchronisation
 * and no allocation).
 */
class StreamGraph {
private:
    gputils_graph_t m_graph = nullptr;
    void *m_stream = nullptr;

public:
    explicit StreamGraph(size_t streamIdx = 0) : m_stream(Session::getInstance().stream(streamIdx)) {}
    ~StreamGraph() {
        if (m_graph) gputils_graph_destroy(m_graph);
    }
    StreamGraph(const StreamGraph &) = delete;
    StreamGraph &operator=(const StreamGraph &) = delete;

    template <typename F>
    void capture(F &&f) {
        if (m_graph) {
            gpuErrChk(gputils_graph_destroy(m_graph));
            m_graph = nullptr;
        }
        gpuErrChk(gputils_graph_capture_begin(m_stream));
        f();
        gpuErrChk(gputils_graph_capture_end(m_stream, &m_graph));
    }

    void launch() const {
        if (!m_graph) throw std::logic_error("[StreamGraph] nothing captured");
        gpuErrChk(gputils_graph_launch(m_graph, m_stream));
    }
};

#endif /* GPUTILS_TENSOR_HPP */
