/*
 * tensor.hpp -- header-only C++ API of GPUtils' batched dense linear
 * algebra (DTensor, Session, IStatus, CholeskyBatchFactoriser, ...) on
 * MI355X.  Source-compatible replacement of GPUtils' include/tensor.cuh:
 * user code that includes this header instead compiles unchanged with any
 * C++17 compiler (no HIP/CUDA compiler needed) and links
 * libgputils_hip.so, whose C ABI (gputils_hip.h) runs the gfx950 kernels.
 *
 * Storage model.  A DTensor is a view (shape, first element, per-matrix
 * pointer table, stream index) plus up to two owned device buffers
 * (DeviceBuffer: one counted allocation, freed with its owner).  A tensor
 * that allocated its elements owns them; slices and operator= targets only
 * view someone else's.  The same holds for the pointer table: an axis-2
 * slice borrows the parent's table (shifted by its first matrix).
 *
 * Deliberate differences from tensor.cuh (each one documented at the
 * method):
 *  - Session exposes HIP-native names (stream(), allocate()) instead of the
 *    vendor-named cuBlasHandle/cuSolverHandle/cudaAllocate.
 *  - Solver calls run on the tensor's stream instead of always stream 0.
 *  - CholeskyBatchFactoriser::solve accepts b with any number of columns.
 *  - Svd::rank() recomputes the rank on each call instead of adding to a
 *    counter that is never reset.
 *  - Svd never modifies the factorised matrix (cuSOLVER's gesvd
 *    overwrites it); Nullspace throws if an SVD did not converge.
 *  - Slices check their range; reshape() keeps the allocated-bytes counter
 *    exact and gives a reshaped slice a pointer table of its new shape.
 *  - operator= releases the storage the target owned before it aliases
 *    (the reference keeps counting it and never frees it), and assigning
 *    from a temporary moves its storage instead of aliasing memory that is
 *    freed at the end of the statement.
 *  - New, with no reference counterpart: MixedCholeskyBatchSolver (FP32
 *    factor + FP64 refinement), RaggedLeastSquares (uneven sizes in one
 *    plan), addAB(..., GemmMode::bf16x3) and StreamGraph (HIP graphs).
 */
#ifndef GPUTILS_TENSOR_HPP
#define GPUTILS_TENSOR_HPP

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <numeric>
#include <optional>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "gputils_hip.h"

#define DEFAULT_FPX double
#define THREADS_PER_BLOCK 512
#define TEMPLATE_WITH_TYPE_T template <typename T = DEFAULT_FPX>
#define TEMPLATE_CONSTRAINT_REQUIRES_FPX

/* ------------------------------------------------------------ helpers --- */

namespace gputils_detail {
/** One process-wide engine for the unseeded helpers below. */
inline std::mt19937_64 &hostEngine() {
    static std::mt19937_64 engine{std::random_device{}()};
    return engine;
}
}  // namespace gputils_detail

/** n uniform reals in [low, hi], unseeded (tensor.cuh:47-58). */
TEMPLATE_WITH_TYPE_T
std::vector<T> generateRealRandomVector(size_t n, T low, T hi) {
    std::uniform_real_distribution<T> draw(low, hi);
    std::vector<T> out;
    out.reserve(n);
    while (out.size() < n) out.push_back(draw(gputils_detail::hostEngine()));
    return out;
}

/** n uniform integers in [low, hi], unseeded (tensor.cuh:60-69). */
inline std::vector<int> generateIntRandomVector(size_t n, int low, int hi) {
    std::uniform_int_distribution<int> draw(low, hi);
    std::vector<int> out;
    out.reserve(n);
    while (out.size() < n) out.push_back(draw(gputils_detail::hostEngine()));
    return out;
}

/** Blocks of threads_per_block threads needed for n items (tensor.cuh:79-81). */
constexpr size_t numBlocks(size_t n, size_t threads_per_block = THREADS_PER_BLOCK) {
    return (n + threads_per_block - 1) / threads_per_block;
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

    /** Number of streams this session holds. */
    size_t numStreams() const { return m_streams.size(); }

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

/* ---------------------------------------------------- device storage --- */

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

/**
 * One device allocation counted in Session's byte total and released by
 * its destructor.  Movable, not copyable.  Copies into and out of device
 * memory go through the null stream after every Session stream has
 * drained, the ordering of the reference's synchronous cudaMemcpy.
 */
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(size_t bytes, bool zero) {
        if (bytes == 0) return;
        gpuErrChk(Session::getInstance().allocate(&m_ptr, bytes));
        m_bytes = bytes;
        if (zero) gpuErrChk(gputils_memset(m_ptr, 0, bytes));
    }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&from) noexcept
        : m_ptr(std::exchange(from.m_ptr, nullptr)), m_bytes(std::exchange(from.m_bytes, 0)) {}
    DeviceBuffer &operator=(DeviceBuffer &&from) noexcept {
        if (this != &from) {
            release();
            m_ptr = std::exchange(from.m_ptr, nullptr);
            m_bytes = std::exchange(from.m_bytes, 0);
        }
        return *this;
    }

    void *get() const { return m_ptr; }
    size_t bytes() const { return m_bytes; }
    bool empty() const { return m_ptr == nullptr; }
    /** Whether address p lies inside this allocation. */
    bool holds(const void *p) const {
        const char *base = static_cast<const char *>(m_ptr), *q = static_cast<const char *>(p);
        return m_ptr != nullptr && q >= base && q < base + m_bytes;
    }

    void release() {
        if (!m_ptr) return;
        gpuErrChk(gputils_free(m_ptr));
        Session::getInstance().incrementAllocatedBytes(-static_cast<long long>(m_bytes));
        m_ptr = nullptr;
        m_bytes = 0;
    }

private:
    void *m_ptr = nullptr;
    size_t m_bytes = 0;
};

/** Blocking copy between host and device memory (kind as gputils_memcpy). */
inline void blockingCopy(void *dst, const void *src, size_t bytes, int kind) {
    if (bytes == 0) return;
    Session::getInstance().synchronizeAllStreams();
    gpuErrChk(gputils_memcpy(dst, src, bytes, kind));
}

/** Rows x cols x mats extent of a tensor. */
struct Shape {
    size_t rows = 0, cols = 0, mats = 0;
    size_t perMatrix() const { return rows * cols; }
    size_t count() const { return rows * cols * mats; }
    bool operator==(const Shape &o) const { return rows == o.rows && cols == o.cols && mats == o.mats; }
    std::string str() const {
        return "(" + std::to_string(rows) + ", " + std::to_string(cols) + ", " + std::to_string(mats) + ")";
    }
};

/** Row-major (per matrix) host data rearranged into column-major order. */
template <typename T>
std::vector<T> rowToColumnMajor(const std::vector<T> &rowMajorData, const Shape &s) {
    std::vector<T> out(rowMajorData.size());
    const size_t per = s.perMatrix();
    for (size_t e = 0; e < out.size(); e++) {
        const size_t mat = e / per, within = e % per;
        const size_t row = within % s.rows, col = within / s.rows;  // e is column-major
        out[e] = rowMajorData[mat * per + row * s.cols + col];
    }
    return out;
}
}  // namespace gputils_detail

/* ------------------------------------------------------------- tensor --- */

/** Precision of a single-precision batched product (no reference counterpart). */
enum class GemmMode {
    standard,  ///< FP32 (or FP64) products and accumulation
    bf16x3     ///< float only: each operand split into two bfloat16 terms, three MFMA products, FP32 accumulation
};

/**
 * (m, n, k) tensor of k column-major m x n matrices on the GPU
 * (tensor.cuh:271-661).  Element (i, j, l) sits at i + m*(j + n*l).
 * Owns its elements unless it is a slice or the target of operator=
 * (which alias).  For k > 1 a device array of per-matrix pointers is kept,
 * as in the reference, so that ptrMatrices() and the byte count match.
 */
TEMPLATE_WITH_TYPE_T
class DTensor {
private:
    using Shape = gputils_detail::Shape;
    using DeviceBuffer = gputils_detail::DeviceBuffer;

    Shape m_shape;
    T *m_data = nullptr;           ///< first element (owned or viewed)
    T **m_table = nullptr;         ///< device table of per-matrix pointers (owned or borrowed)
    DeviceBuffer m_elements;       ///< non-empty iff this tensor owns its elements
    DeviceBuffer m_tableStorage;   ///< non-empty iff this tensor owns its pointer table
    size_t m_idxStream = 0;

    /** Allocates the elements (and, for several matrices, the pointer table). */
    void claimStorage(bool zero) {
        if (m_shape.count() == 0) return;
        m_elements = DeviceBuffer(m_shape.count() * sizeof(T), zero);
        m_data = static_cast<T *>(m_elements.get());
        rebuildPointerTable();
    }

    /**
     * Makes the pointer table describe the current shape: for more than one
     * matrix the table is owned (reused when large enough, else replaced)
     * and refilled by a kernel; a borrowed table is kept only while the
     * matrix size and count it was built for still hold.
     */
    void rebuildPointerTable(bool borrowedStillValid = false) {
        if (m_shape.mats <= 1 || borrowedStillValid) return;
        const size_t need = m_shape.mats * sizeof(T *);
        if (m_tableStorage.bytes() < need) {
            m_tableStorage = DeviceBuffer(need, false);
            m_table = static_cast<T **>(m_tableStorage.get());
        } else if (m_table != m_tableStorage.get()) {
            m_table = static_cast<T **>(m_tableStorage.get());
        }
        gpuErrChk(gputils_fill_pointer_array(reinterpret_cast<void **>(m_table), m_data,
                                             static_cast<int64_t>(m_shape.perMatrix() * sizeof(T)),
                                             static_cast<int64_t>(m_shape.mats), stream()));
        gpuErrChk(gputils_stream_synchronize(stream()));
    }

    std::ostream &print(std::ostream &out) const {
        std::vector<T> host;
        download(host);
        out << "DTensor" << m_shape.str() << ", column-major\n";
        for (size_t l = 0; l < m_shape.mats; l++) {
            out << "  matrix " << l << ":\n";
            for (size_t i = 0; i < m_shape.rows; i++) {
                out << "   ";
                for (size_t j = 0; j < m_shape.cols; j++)
                    out << ' ' << std::setw(12) << host[i + m_shape.rows * (j + m_shape.cols * l)];
                out << '\n';
            }
        }
        return out;
    }

    /** this (a fresh copy of `base`) += sign * other; for operator+ and -. */
    static DTensor combined(const DTensor &base, const DTensor &other, T sign) {
        gputils_detail::require_fp<T>();
        DTensor out(base);
        out.axpyFrom(sign, other);
        return out;
    }

    void axpyFrom(T alpha, const DTensor &x) {
        if constexpr (std::is_same_v<T, double>) gpuErrChk(gputils_daxpy(numEl(), alpha, x.m_data, m_data, stream()));
        if constexpr (std::is_same_v<T, float>) gpuErrChk(gputils_saxpy(numEl(), alpha, x.m_data, m_data, stream()));
    }

    template <typename S>
    S reduceToHost(int (*fn)(int64_t, const S *, S *, void *, void *)) const;

public:
    /** Stream handle used by operations on this tensor. */
    void *stream() const { return Session::getInstance().stream(m_idxStream); }

    /**
     * Selects the stream that operations on this tensor use; returns a copy
     * of the tensor by value, as the reference does (tensor.cuh:663-670).
     */
    DTensor<T> setStreamIdx(size_t idx) {
        if (idx >= s_numStreams) throw std::invalid_argument("Invalid stream index; it exceeds the max allocated streams");
        m_idxStream = idx;
        return *this;
    }

    size_t streamIdx() const { return m_idxStream; }

    static DTensor<T> createRandomTensor(size_t numRows, size_t numCols, size_t numMats, T low, T hi);
    static DTensor<T> parseFromFile(std::string path_to_file, StorageMode mode = StorageMode::defaultMajor);

    DTensor() = default;
    ~DTensor() = default;

    /** (m,n,k) tensor, optionally zeroed (tensor.cuh:863-872). */
    DTensor(size_t m, size_t n = 1, size_t k = 1, bool zero = false) : m_shape{m, n, k} { claimStorage(zero); }

    /** (m,n,k) tensor holding `data`, given in column- or row-major order (tensor.cuh:874-884). */
    DTensor(const std::vector<T> &data, size_t m, size_t n = 1, size_t k = 1,
            StorageMode mode = StorageMode::defaultMajor)
        : m_shape{m, n, k} {
        claimStorage(false);
        upload(data, mode);
    }

    /** Deep copy on the device (tensor.cuh:886-897). */
    DTensor(const DTensor &other) : m_shape(other.m_shape), m_idxStream(other.m_idxStream) {
        claimStorage(false);
        gputils_detail::blockingCopy(m_data, other.m_data, numEl() * sizeof(T), 2);
    }

    /** Takes over other's storage and views; other is left empty (tensor.cuh:927-946). */
    DTensor(DTensor &&other) noexcept
        : m_shape(std::exchange(other.m_shape, Shape{})),
          m_data(std::exchange(other.m_data, nullptr)),
          m_table(std::exchange(other.m_table, nullptr)),
          m_elements(std::move(other.m_elements)),
          m_tableStorage(std::move(other.m_tableStorage)),
          m_idxStream(other.m_idxStream) {}

    /**
     * View of part of `parent` (tensor.cuh:899-925), `to` inclusive.
     * axis 2: matrices from..to, borrowing the parent's pointer table;
     * axis 1: columns from..to of the first matrix; axis 0: elements
     * from..to of the flat element array, as a column vector.  Ranges past
     * the parent's extent throw (the reference does not check).
     */
    DTensor(const DTensor &parent, size_t axis, size_t from, size_t to) : m_idxStream(parent.m_idxStream) {
        if (from > to) throw std::invalid_argument("from > to");
        const size_t span = to - from + 1;
        const Shape &p = parent.m_shape;
        switch (axis) {
            case 2:
                if (to >= p.mats) throw std::invalid_argument("[slice] matrix index out of range");
                m_shape = Shape{p.rows, p.cols, span};
                m_data = parent.m_data + from * p.perMatrix();
                m_table = parent.m_table ? parent.m_table + from : nullptr;
                break;
            case 1:
                if (to >= p.cols) throw std::invalid_argument("[slice] column index out of range");
                m_shape = Shape{p.rows, span, 1};
                m_data = parent.m_data + from * p.rows;
                break;
            case 0:
                if (to >= p.count()) throw std::invalid_argument("[slice] element index out of range");
                m_shape = Shape{span, 1, 1};
                m_data = parent.m_data + from;
                break;
            default:
                throw std::invalid_argument("[slice] axis must be 0, 1 or 2");
        }
    }

    T *raw() const { return m_data; }
    T **ptrMatrices() const { return m_table; }
    size_t numRows() const { return m_shape.rows; }
    size_t numCols() const { return m_shape.cols; }
    size_t numMats() const { return m_shape.mats; }
    size_t numEl() const { return m_shape.count(); }

    /** Blocking upload of numEl() values in column- or row-major order (tensor.cuh:1128-1145). */
    bool upload(const std::vector<T> &vec, StorageMode mode = StorageMode::defaultMajor) {
        if (vec.size() != numEl()) throw std::invalid_argument("[upload] vec has wrong size");
        if (mode == StorageMode::rowMajor) {
            const std::vector<T> columnMajorData = gputils_detail::rowToColumnMajor(vec, m_shape);
            gputils_detail::blockingCopy(m_data, columnMajorData.data(), numEl() * sizeof(T), 0);
        } else {
            gputils_detail::blockingCopy(m_data, vec.data(), numEl() * sizeof(T), 0);
        }
        return true;
    }

    /** Blocking download of all elements, column-major (tensor.cuh:1147-1154). */
    void download(std::vector<T> &vec) const {
        vec.resize(numEl());
        gputils_detail::blockingCopy(vec.data(), m_data, numEl() * sizeof(T), 1);
    }

    /** Device-to-device copy into a tensor with at least numEl() elements (tensor.cuh:1199-1208). */
    void deviceCopyTo(DTensor<T> &elsewhere) const {
        if (elsewhere.numEl() < numEl()) throw std::invalid_argument("[deviceCopyTo] tensor does not fit into destination");
        gputils_detail::blockingCopy(elsewhere.raw(), m_data, numEl() * sizeof(T), 2);
    }

    /** Rows rowsFrom..rowsTo (inclusive) of matrix matIdx as a new tensor (tensor.cuh:1396-1424). */
    DTensor<T> getRows(size_t rowsFrom, size_t rowsTo, size_t matIdx) const {
        if (rowsFrom > rowsTo || rowsTo >= m_shape.rows || matIdx >= m_shape.mats)
            throw std::invalid_argument("[getRows] index out of range");
        const size_t len = rowsTo - rowsFrom + 1;
        DTensor<T> rowsOnly(len, m_shape.cols, 1);
        const T *src = m_data + matIdx * m_shape.perMatrix() + rowsFrom;
        // rows of a column-major matrix: one contiguous copy per column
        Session::getInstance().synchronizeAllStreams();
        for (size_t c = 0; c < m_shape.cols; c++)
            gpuErrChk(gputils_memcpy(rowsOnly.raw() + c * len, src + c * m_shape.rows, len * sizeof(T), 2));
        return rowsOnly;
    }

    /** Batched transpose (tensor.cuh:1167-1197, one launch for all matrices). */
    DTensor<T> tr() const {
        gputils_detail::require_fp<T>();
        DTensor<T> out(m_shape.cols, m_shape.rows, m_shape.mats);
        const int64_t m = (int64_t)m_shape.rows, n = (int64_t)m_shape.cols, per = m * n;
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dtranspose_batched(m, n, raw(), m, per, out.raw(), n, per, (int64_t)m_shape.mats, stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_stranspose_batched(m, n, raw(), m, per, out.raw(), n, per, (int64_t)m_shape.mats, stream()));
        return out;
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
    void addAB(const DTensor<T> &A, const DTensor<T> &B, T alpha = 1, T beta = 0) {
        addAB(A, B, alpha, beta, GemmMode::standard);
    }

    /**
     * As addAB above with a choice of arithmetic (no reference
     * counterpart): GemmMode::bf16x3 (float only) runs three bfloat16 MFMA
     * products of the split operands with FP32 accumulation.
     */
    void addAB(const DTensor<T> &A, const DTensor<T> &B, T alpha, T beta, GemmMode mode);

    /**
     * Reinterprets the elements as (newNumRows, newNumCols, newNumMats)
     * (tensor.cuh:828-861); the element count must not change.  An owned
     * pointer table is refilled (and reallocated only when it is too small);
     * a slice's borrowed table is replaced by an owned one unless the matrix
     * size is unchanged and the matrix count does not grow.  The byte
     * counter always changes by exactly what is allocated or freed.
     */
    void reshape(size_t newNumRows, size_t newNumCols, size_t newNumMats = 1) {
        const Shape next{newNumRows, newNumCols, newNumMats};
        if (next == m_shape) return;
        if (next.count() != m_shape.count())
            throw std::invalid_argument("[reshape] cannot view the " + std::to_string(m_shape.count()) +
                                        " elements of DTensor" + m_shape.str() + " as DTensor" + next.str() +
                                        " (" + std::to_string(next.count()) + " elements)");
        const bool borrowed = m_table != nullptr && m_tableStorage.empty();
        const bool borrowedStillValid = borrowed && next.perMatrix() == m_shape.perMatrix() && next.mats <= m_shape.mats;
        m_shape = next;
        rebuildPointerTable(borrowedStillValid);
    }

    /** Save as text, or as binary when the name ends in .bt (tensor.cuh:797-825). */
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

    /**
     * Alias: this tensor views other's elements and pointer table
     * (tensor.cuh:1219-1228).  Storage this tensor owned is released first.
     */
    DTensor &operator=(const DTensor &other) {
        if (this == &other) return *this;
        // a view into this tensor's own storage keeps that storage alive
        if (!m_elements.holds(other.m_data)) m_elements.release();
        if (!m_tableStorage.holds(other.m_table)) m_tableStorage.release();
        m_shape = other.m_shape;
        m_data = other.m_data;
        m_table = other.m_table;
        m_idxStream = other.m_idxStream;
        return *this;
    }

    /** Takes over a temporary's storage (an alias of it would dangle). */
    DTensor &operator=(DTensor &&other) noexcept {
        if (this == &other) return *this;
        if (other.m_elements.empty()) return *this = static_cast<const DTensor &>(other);  // a view: alias it
        m_elements = std::move(other.m_elements);
        m_tableStorage = std::move(other.m_tableStorage);
        m_shape = std::exchange(other.m_shape, Shape{});
        m_data = std::exchange(other.m_data, nullptr);
        m_table = std::exchange(other.m_table, nullptr);
        m_idxStream = other.m_idxStream;
        return *this;
    }

    /** Element (i,j,k) by a blocking single-element copy (tensor.cuh:1278-1284). */
    T operator()(size_t i, size_t j = 0, size_t k = 0) const {
        const size_t at = i + m_shape.rows * (j + m_shape.cols * k);
        if (at >= numEl()) throw std::out_of_range("[DTensor] element index out of range");
        T value;
        gputils_detail::blockingCopy(&value, m_data + at, sizeof(T), 1);
        return value;
    }

    DTensor &operator*=(T scalar) {
        gputils_detail::require_fp<T>();
        if constexpr (std::is_same_v<T, double>) gpuErrChk(gputils_dscal(numEl(), scalar, m_data, stream()));
        if constexpr (std::is_same_v<T, float>) gpuErrChk(gputils_sscal(numEl(), scalar, m_data, stream()));
        return *this;
    }

    DTensor &operator+=(const DTensor &rhs) {
        gputils_detail::require_fp<T>();
        axpyFrom(T(1), rhs);
        return *this;
    }

    DTensor &operator-=(const DTensor &rhs) {
        gputils_detail::require_fp<T>();
        axpyFrom(T(-1), rhs);
        return *this;
    }

    friend DTensor operator+(DTensor &first, const DTensor &second) { return combined(first, second, T(1)); }

    friend DTensor operator-(DTensor &first, const DTensor &second) { return combined(first, second, T(-1)); }

    /** New tensor A_i B_i with B's matrix count (tensor.cuh:645-650). */
    friend DTensor<T> operator*(DTensor &A, DTensor &B) {
        DTensor<T> product(A.numRows(), B.numCols(), B.numMats());
        product.addAB(A, B);
        return product;
    }

    friend DTensor<T> operator*(T a, DTensor &B) {
        DTensor<T> scaled(B);
        scaled *= a;
        return scaled;
    }

    friend std::ostream &operator<<(std::ostream &out, const DTensor<T> &data) { return data.print(out); }
};

/* --- reductions: one device result copied back on the tensor's stream --- */

template <typename T>
template <typename S>
S DTensor<T>::reduceToHost(int (*fn)(int64_t, const S *, S *, void *, void *)) const {
    Session &s = Session::getInstance();
    S *slot = static_cast<S *>(s.scalarSlot(m_idxStream));
    gpuErrChk(fn((int64_t)numEl(), m_data, slot, s.reduceWorkspace(m_idxStream), stream()));
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
    if (!(m_shape == other.m_shape)) throw std::invalid_argument("[dotF] incompatible dimensions");
    Session &s = Session::getInstance();
    T *slot = static_cast<T *>(s.scalarSlot(m_idxStream));
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_ddot(numEl(), m_data, other.m_data, slot, s.reduceWorkspace(m_idxStream), stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_sdot(numEl(), m_data, other.m_data, slot, s.reduceWorkspace(m_idxStream), stream()));
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
void DTensor<T>::addAB(const DTensor<T> &A, const DTensor<T> &B, T alpha, T beta, GemmMode mode) {
    gputils_detail::require_fp<T>();
    const int64_t m = (int64_t)A.numRows(), inner = (int64_t)A.numCols(), n = (int64_t)B.numCols();
    const int64_t count = (int64_t)A.numMats();
    /* like the reference: sizes from A and B, k from A, no further checks */
    if constexpr (std::is_same_v<T, double>) {
        if (mode != GemmMode::standard) throw std::invalid_argument("[addAB] GemmMode::bf16x3 needs T = float");
        gpuErrChk(gputils_dgemm_strided_batched(m, n, inner, alpha, A.raw(), m, m * inner, B.raw(), inner, inner * n,
                                                beta, raw(), m, m * n, count, stream()));
    }
    if constexpr (std::is_same_v<T, float>) {
        const int abiMode = mode == GemmMode::bf16x3 ? GPUTILS_GEMM_BF16X3 : GPUTILS_GEMM_FP32;
        gpuErrChk(gputils_sgemm_strided_batched(m, n, inner, alpha, A.raw(), m, m * inner, B.raw(), inner, inner * n,
                                                beta, raw(), m, m * n, count, abiMode, stream()));
    }
}

/**
 * Batched least squares (tensor.cuh:1340-1394): for every i, b_i[0:n] <-
 * argmin ||A_i x - b_i|| (m >= n, one right-hand side).  A_i is overwritten
 * with its Householder QR factors.
 */
template <typename T>
void DTensor<T>::leastSquaresBatched(DTensor &B) {
    gputils_detail::require_fp<T>();
    if (B.numRows() != numRows())
        throw std::invalid_argument("[Least squares batched] rhs rows does not equal lhs rows");
    if (B.numCols() != 1) throw std::invalid_argument("[Least squares batched] rhs are not vectors");
    if (B.numMats() != numMats())
        throw std::invalid_argument("[Least squares batched] rhs numMats does not equal lhs numMats");
    if (numCols() > numRows())
        throw std::invalid_argument("[Least squares batched] supports square or tall matrices only");
    const int64_t m = (int64_t)numRows(), n = (int64_t)numCols(), count = (int64_t)numMats();
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_dgels_batched(m, n, 1, raw(), m, m * n, B.raw(), m, m, nullptr, count, stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_sgels_batched(m, n, 1, raw(), m, m * n, B.raw(), m, m, nullptr, count, stream()));
}

/** Tensor of uniform random entries in [low, hi]; unseeded (tensor.cuh:690-702). */
template <typename T>
DTensor<T> DTensor<T>::createRandomTensor(size_t numRows, size_t numCols, size_t numMats, T low, T hi) {
    const size_t count = numRows * numCols * numMats;
    if constexpr (std::is_floating_point_v<T>) {
        return DTensor<T>(generateRealRandomVector<T>(count, low, hi), numRows, numCols, numMats);
    } else if constexpr (std::is_same_v<T, int>) {
        return DTensor<T>(generateIntRandomVector(count, low, hi), numRows, numCols, numMats);
    } else {
        throw std::invalid_argument("[createRandomTensor] unsupported type T");
    }
}

/* --------------------------------------------------------- file I/O --- */
/*
 * Two on-disk forms of an (m, n, k) tensor (tensor.cuh:704-825):
 *  - text: m, n and k on the first three lines, then the m*n*k elements in
 *    column-major order, one per line (missing trailing values read as 0);
 *  - binary (file name ending in ".bt"): m, n, k as three little-endian
 *    uint64 values, then the elements as raw sizeof(T)-byte values in
 *    column-major order.
 */
namespace gputils_detail {

inline bool isBinaryTensorFile(const std::string &path) {
    static const std::string ext = ".bt";
    return path.size() >= ext.size() && path.compare(path.size() - ext.size(), ext.size(), ext) == 0;
}

/** One text field as an arithmetic value of type V. */
template <typename V>
V parseTextValue(const std::string &field) {
    static_assert(std::is_arithmetic_v<V>, "tensor files hold arithmetic element types only");
    const char *begin = field.c_str();
    char *end = nullptr;
    errno = 0;
    V value;
    if constexpr (std::is_floating_point_v<V>) {
        value = static_cast<V>(std::strtold(begin, &end));
    } else if constexpr (std::is_signed_v<V>) {
        value = static_cast<V>(std::strtoll(begin, &end, 10));
    } else {
        value = static_cast<V>(std::strtoull(begin, &end, 10));
    }
    if (end == begin) throw std::invalid_argument("[parseFromFile] not a number: '" + field + "'");
    return value;
}

template <typename T>
struct TensorFileContents {
    Shape shape;
    std::vector<T> values;
};

template <typename T>
TensorFileContents<T> readTextTensor(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::invalid_argument("[parseFromFile] cannot open '" + path + "'");
    size_t dims[3];
    std::string line;
    for (size_t &d : dims) {
        if (!std::getline(in, line)) throw std::invalid_argument("[parseFromFile] '" + path + "' has no header");
        d = parseTextValue<unsigned long long>(line);
    }
    TensorFileContents<T> f{Shape{dims[0], dims[1], dims[2]}, {}};
    f.values.assign(f.shape.count(), T(0));
    for (T &v : f.values) {
        if (!std::getline(in, line)) break;
        v = parseTextValue<T>(line);
    }
    return f;
}

template <typename T>
TensorFileContents<T> readBinaryTensor(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::invalid_argument("[parseFromFile] cannot open '" + path + "'");
    uint64_t dims[3];
    if (!in.read(reinterpret_cast<char *>(dims), sizeof dims))
        throw std::invalid_argument("[parseFromFile] '" + path + "' is shorter than its header");
    TensorFileContents<T> f{Shape{dims[0], dims[1], dims[2]}, {}};
    f.values.resize(f.shape.count());
    const std::streamsize payload = static_cast<std::streamsize>(f.values.size() * sizeof(T));
    in.read(reinterpret_cast<char *>(f.values.data()), payload);
    if (in.gcount() != payload) throw std::invalid_argument("[parseFromFile] '" + path + "' is truncated");
    return f;
}

template <typename T>
void writeTensorFile(const std::string &path, const Shape &shape, const std::vector<T> &values) {
    if (isBinaryTensorFile(path)) {
        std::ofstream out(path, std::ios::binary);
        const uint64_t dims[3] = {shape.rows, shape.cols, shape.mats};
        out.write(reinterpret_cast<const char *>(dims), sizeof dims);
        out.write(reinterpret_cast<const char *>(values.data()), (std::streamsize)(values.size() * sizeof(T)));
        if (!out) throw std::runtime_error("[saveToFile] cannot write '" + path + "'");
        return;
    }
    std::ostringstream text;
    if constexpr (std::is_floating_point_v<T>) text << std::setprecision(std::numeric_limits<T>::max_digits10);
    text << shape.rows << '\n' << shape.cols << '\n' << shape.mats << '\n';
    for (const T &v : values) text << v << '\n';
    std::ofstream out(path);
    out << text.str();
    if (!out) throw std::runtime_error("[saveToFile] cannot write '" + path + "'");
}
}  // namespace gputils_detail

template <typename T>
DTensor<T> DTensor<T>::parseFromFile(std::string path_to_file, StorageMode mode) {
    auto contents = gputils_detail::isBinaryTensorFile(path_to_file) ? gputils_detail::readBinaryTensor<T>(path_to_file)
                                                                     : gputils_detail::readTextTensor<T>(path_to_file);
    const auto &s = contents.shape;
    return DTensor<T>(contents.values, s.rows, s.cols, s.mats, mode);
}

template <typename T>
void DTensor<T>::saveToFile(std::string pathToFile) {
    std::vector<T> host;
    download(host);
    gputils_detail::writeTensorFile(pathToFile, m_shape, host);
}

template <typename T>
void DTensor<T>::applyRightGivensRotation(size_t i, size_t j, const T *c, const T *minus_s) {
    gputils_detail::require_fp<T>();
    if (numMats() > 1) throw std::invalid_argument("[applyRightGivensRotation] tensors (nMat>1) not supported");
    if (i >= numCols() || j >= numCols()) throw std::invalid_argument("[applyRightGivensRotation] invalid column index");
    T *col_i = m_data + i * numRows();
    T *col_j = m_data + j * numRows();
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_drot((int64_t)numRows(), col_i, 1, col_j, 1, *c, *minus_s, stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_srot((int64_t)numRows(), col_i, 1, col_j, 1, *c, *minus_s, stream()));
}

template <typename T>
void DTensor<T>::applyLeftGivensRotation(size_t i, size_t j, const T *c, const T *minus_s) {
    gputils_detail::require_fp<T>();
    if (numMats() > 1) throw std::invalid_argument("[applyLeftGivensRotation] tensors (nMat>1) not supported");
    if (i >= numRows() || j >= numRows()) throw std::invalid_argument("[applyLeftGivensRotation] invalid row index");
    const int64_t ld = (int64_t)numRows();
    if constexpr (std::is_same_v<T, double>)
        gpuErrChk(gputils_drot((int64_t)numCols(), m_data + i, ld, m_data + j, ld, *c, *minus_s, stream()));
    if constexpr (std::is_same_v<T, float>)
        gpuErrChk(gputils_srot((int64_t)numCols(), m_data + i, ld, m_data + j, ld, *c, *minus_s, stream()));
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
 * Mixed-precision batched SPD solve (no reference counterpart; it sits
 * beside CholeskyBatchFactoriser, tensor.cuh:2098-2197).  Each A_i
 * (FP64, lower triangle used, read only) is factorised in FP32 on the GPU;
 * solves are refined in FP64 until every right-hand side meets
 *   ||b - A x||_inf <= tol (||A||_inf ||x||_inf + ||b||_inf)
 * (LAPACK DSPOSV's test).  Systems that do not converge in maxIter steps,
 * or whose FP32 factorisation fails, are solved again in FP64: then
 * iterations()(0,0,i) = -1 and info()(0,0,i) is the FP64 status (0, or
 * the order of the first non-positive leading minor).  A must outlive this
 * object.  n <= 128 runs the fused mixed kernel; larger n uses the FP64
 * kernels directly (iterations = -1).
 */
class MixedCholeskyBatchSolver : public IStatus {
private:
    DTensor<double> *m_matrix;
    DTensor<int> m_iters;
    gputils_detail::DeviceBuffer m_work;  ///< FP64 copy of A, only for n beyond the fused kernel
    double m_tol;
    int m_maxIter;

public:
    MixedCholeskyBatchSolver(DTensor<double> &A, double tol = 1e-12, int maxIter = 30)
        : IStatus(A.numMats()), m_matrix(&A), m_iters(1, 1, A.numMats(), true), m_tol(tol), m_maxIter(maxIter) {
        if (A.numRows() != A.numCols()) throw std::invalid_argument("[MixedCholesky] A must be square");
        if (!(tol > 0)) throw std::invalid_argument("[MixedCholesky] tol must be positive");
        if (maxIter < 0) throw std::invalid_argument("[MixedCholesky] maxIter must be non-negative");
        const int64_t wb = gputils_dposv_mixed_workspace_bytes((int64_t)A.numRows(), (int64_t)A.numMats());
        if (wb < 0) throw std::invalid_argument("[MixedCholesky] unsupported size");
        m_work = gputils_detail::DeviceBuffer((size_t)wb, false);
    }

    /** x_i <- A_i^{-1} b_i for every column of b_i; b is read only, x has b's shape. */
    void solve(const DTensor<double> &b, DTensor<double> &x) {
        const size_t n = m_matrix->numRows(), k = m_matrix->numMats();
        if (b.numRows() != n || b.numMats() != k || b.numCols() < 1)
            throw std::invalid_argument("[MixedCholesky] A and b incompatible");
        if (x.numRows() != n || x.numCols() != b.numCols() || x.numMats() != k)
            throw std::invalid_argument("[MixedCholesky] x must have the shape of b");
        if (x.raw() == b.raw()) throw std::invalid_argument("[MixedCholesky] x must not alias b");
        const int64_t nn = (int64_t)n, nrhs = (int64_t)b.numCols();
        gpuErrChk(gputils_dposv_mixed_batched(nn, nrhs, m_matrix->raw(), nn, nn * nn, b.raw(), nn, nn * nrhs, x.raw(), nn,
                                              nn * nrhs, m_tol, m_maxIter, m_iters.raw(), m_info->raw(), m_work.get(),
                                              (int64_t)k, m_matrix->stream()));
    }

    /** Refinement steps used per system (-1: solved by the FP64 fallback). */
    DTensor<int> &iterations() { return m_iters; }
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

/**
 * Least squares over many systems of uneven sizes in one plan (no
 * reference counterpart: leastSquaresBatched, tensor.cuh:1340-1394, needs
 * one size for all).  System s is rows[s] x cols[s] (1 <= cols <= rows
 * <= 512, cols <= 64), stored column-major at element offsetA(s) of a flat
 * FP64 tensor A, with its right-hand side at offsetB(s) of a flat tensor b;
 * the default layout packs the systems back to back.  solve() overwrites
 * the first cols[s] entries of each right-hand side with the solution of
 * an FP32 Householder QR (FP64 norms) refined refineSteps times in FP64;
 * A is read only.  tune() times the launch shapes of every size bucket on
 * the caller's data and keeps the fastest (plan-time work: it allocates
 * and synchronises).
 */
class RaggedLeastSquares {
private:
    gputils_gels_plan_t m_plan = nullptr;
    std::vector<int64_t> m_offA, m_offB;
    size_t m_elementsA = 0, m_elementsB = 0;

    void checkData(const DTensor<double> &A, const DTensor<double> &b) const {
        if (A.numEl() < m_elementsA) throw std::invalid_argument("[RaggedLeastSquares] A holds too few elements");
        if (b.numEl() < m_elementsB) throw std::invalid_argument("[RaggedLeastSquares] b holds too few elements");
    }

public:
    RaggedLeastSquares(const std::vector<int32_t> &rows, const std::vector<int32_t> &cols) {
        if (rows.size() != cols.size()) throw std::invalid_argument("[RaggedLeastSquares] rows and cols differ in length");
        m_offA.resize(rows.size());
        m_offB.resize(rows.size());
        for (size_t s = 0; s < rows.size(); s++) {
            if (rows[s] < 1 || cols[s] < 1) throw std::invalid_argument("[RaggedLeastSquares] empty system");
            m_offA[s] = (int64_t)m_elementsA;
            m_offB[s] = (int64_t)m_elementsB;
            m_elementsA += (size_t)rows[s] * (size_t)cols[s];
            m_elementsB += (size_t)rows[s];
        }
        const int st = gputils_gels_ragged_plan_create(&m_plan, (int64_t)rows.size(), rows.data(), cols.data(),
                                                       m_offA.data(), m_offB.data());
        if (st < 0) throw std::invalid_argument(std::string("[RaggedLeastSquares] ") + gputils_last_error());
        gpuErrChk(st);
    }
    ~RaggedLeastSquares() {
        if (m_plan) gputils_gels_ragged_plan_destroy(m_plan);
    }
    RaggedLeastSquares(const RaggedLeastSquares &) = delete;
    RaggedLeastSquares &operator=(const RaggedLeastSquares &) = delete;

    size_t numSystems() const { return m_offA.size(); }
    /** Elements of the packed A and b tensors this plan addresses. */
    size_t elementsA() const { return m_elementsA; }
    size_t elementsB() const { return m_elementsB; }
    int64_t offsetA(size_t s) const { return m_offA.at(s); }
    int64_t offsetB(size_t s) const { return m_offB.at(s); }

    void tune(const DTensor<double> &A, const DTensor<double> &b, int refineSteps = 1) {
        checkData(A, b);
        gpuErrChk(gputils_gels_ragged_plan_tune(m_plan, A.raw(), b.raw(), refineSteps, A.stream()));
    }

    void solve(const DTensor<double> &A, DTensor<double> &b, int refineSteps = 2) {
        checkData(A, b);
        if (refineSteps < 0) throw std::invalid_argument("[RaggedLeastSquares] refineSteps must be non-negative");
        gpuErrChk(gputils_dgels_ragged(m_plan, A.raw(), b.raw(), refineSteps, A.stream()));
    }
};

/* --------------------------------------------------------------- SVD --- */

/**
 * Batched SVD of tall matrices, A_i = U_i S_i V_i' (tensor.cuh:1500-1676).
 * All matrices are factorised by one launch (the reference loops over
 * them on the host).  rank() counts sigma > epsilon afresh on every call
 * (the reference adds to a counter that is never reset).  A is not
 * modified by factorise(); with destroyMatrix = false a private copy is
 * factorised, as in the reference.
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
    gputils_detail::DeviceBuffer m_work;
    bool m_computeU = false;

public:
    Svd(DTensor<T> &mat, bool computeU = false, bool destroyMatrix = true) : IStatus(mat.numMats()) {
        gputils_detail::require_fp<T>();
        if (mat.numRows() < mat.numCols()) throw std::invalid_argument("[svd] your matrix is fat (no offence)");
        const size_t m = mat.numRows(), n = mat.numCols(), k = mat.numMats();
        const int64_t wb = gputils_gesvd_workspace_bytes((int)m, (int)n, (int64_t)k);
        if (wb < 0) throw std::invalid_argument("[svd] unsupported matrix size");
        if (destroyMatrix) {
            m_tensor = &mat;
        } else {
            m_copy = std::make_unique<DTensor<T>>(mat);
            m_tensor = m_copy.get();
        }
        m_computeU = computeU;
        m_Vtr = std::make_shared<DTensor<T>>(n, n, k);
        m_S = std::make_shared<DTensor<T>>(std::min(m, n), 1, k);
        m_rank = std::make_shared<DTensor<unsigned int>>(1, 1, k, true);
        if (computeU) m_U = std::make_shared<DTensor<T>>(m, m, k);
        m_work = gputils_detail::DeviceBuffer((size_t)wb, false);
    }

    Svd(const Svd &) = delete;
    Svd &operator=(const Svd &) = delete;

    void factorise() {
        const int m = (int)m_tensor->numRows(), n = (int)m_tensor->numCols();
        const int64_t k = (int64_t)m_tensor->numMats(), mn = (int64_t)m * n;
        T *U = m_computeU ? m_U->raw() : nullptr;
        if constexpr (std::is_same_v<T, double>)
            gpuErrChk(gputils_dgesvd_batched(m, n, m_tensor->raw(), m, mn, m_S->raw(), n, U, m, (int64_t)m * m, m_Vtr->raw(),
                                             n, (int64_t)n * n, m_work.get(), m_info->raw(), nullptr, k, m_tensor->stream()));
        if constexpr (std::is_same_v<T, float>)
            gpuErrChk(gputils_sgesvd_batched(m, n, m_tensor->raw(), m, mn, m_S->raw(), n, U, m, (int64_t)m * m, m_Vtr->raw(),
                                             n, (int64_t)n * n, m_work.get(), m_info->raw(), nullptr, k, m_tensor->stream()));
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
 * over the matrices on the host).  Throws std::runtime_error if the SVD of
 * any matrix did not converge.
 */
TEMPLATE_WITH_TYPE_T
class Nullspace {
private:
    std::unique_ptr<DTensor<T>> m_nullspace;
    std::unique_ptr<DTensor<T>> m_projOp;
    std::unique_ptr<DTensor<unsigned int>> m_rank;

public:
    Nullspace(DTensor<T> &a) {
        gputils_detail::require_fp<T>();
        const size_t m = a.numRows(), n = a.numCols(), nMats = a.numMats();
        if (m > n) throw std::invalid_argument("[nullspace] I was expecting a square or fat matrix");
        const int64_t nn = (int64_t)n * n, k = (int64_t)nMats;
        const int64_t wb = gputils_gesvd_workspace_bytes((int)n, (int)m, k);
        if (wb < 0) throw std::invalid_argument("[nullspace] unsupported matrix size");
        m_nullspace = std::make_unique<DTensor<T>>(n, n, nMats, true);
        m_projOp = std::make_unique<DTensor<T>>(n, n, nMats, true);
        DTensor<T> aTranspose = a.tr();
        // SVD of a' (n x m) without right singular vectors: the sweeps run on
        // R^T and U comes from the accumulated rotations (same subspaces)
        DTensor<T> S(m, 1, nMats), U(n, n, nMats);
        DTensor<int> info(1, 1, nMats);
        m_rank = std::make_unique<DTensor<unsigned int>>(1, 1, nMats, true);
        DTensor<unsigned int> &rank = *m_rank;
        gputils_detail::DeviceBuffer work((size_t)wb, false);
        void *st = aTranspose.stream();
        if constexpr (std::is_same_v<T, double>) {
            gpuErrChk(gputils_dgesvd_batched((int)n, (int)m, aTranspose.raw(), n, (int64_t)n * m, S.raw(), m, U.raw(),
                                             n, nn, nullptr, m, (int64_t)m * m, work.get(), info.raw(), nullptr, k, st));
            gpuErrChk(gputils_dsvd_rank_batched((int)m, S.raw(), m, 1e-6, rank.raw(), k, st));
            gpuErrChk(gputils_dnullspace_build_batched((int)n, U.raw(), n, nn, rank.raw(), m_nullspace->raw(), n, nn,
                                                       m_projOp->raw(), n, nn, k, st));
        }
        if constexpr (std::is_same_v<T, float>) {
            gpuErrChk(gputils_sgesvd_batched((int)n, (int)m, aTranspose.raw(), n, (int64_t)n * m, S.raw(), m, U.raw(),
                                             n, nn, nullptr, m, (int64_t)m * m, work.get(), info.raw(), nullptr, k, st));
            gpuErrChk(gputils_ssvd_rank_batched((int)m, S.raw(), m, 1e-6f, rank.raw(), k, st));
            gpuErrChk(gputils_snullspace_build_batched((int)n, U.raw(), n, nn, rank.raw(), m_nullspace->raw(), n, nn,
                                                       m_projOp->raw(), n, nn, k, st));
        }
        gpuErrChk(gputils_stream_synchronize(st));
        std::vector<int> status;
        info.download(status);
        for (size_t i = 0; i < status.size(); i++)
            if (status[i] != 0)
                throw std::runtime_error("[nullspace] the SVD of matrix " + std::to_string(i) + " did not converge");
    }

    DTensor<T> const &nullspace() const { return *m_nullspace; }

    /** b_i <- P_i b_i in place (tensor.cuh:2083-2085).  A single column per
     * matrix is projected as N_i (N_i' b_i) from the nullity columns of N_i
     * (fewer bytes read than P_i); several columns go through P. */
    void project(DTensor<T> &b) {
        const size_t n = m_nullspace->numRows();
        if (b.numRows() == n && b.numCols() == 1 && b.numMats() == m_nullspace->numMats()) {
            const int64_t nn = (int64_t)n * n, k = (int64_t)b.numMats();
            if constexpr (std::is_same_v<T, double>) {
                gpuErrChk(gputils_dnullspace_project_batched((int)n, m_nullspace->raw(), n, nn, m_rank->raw(),
                                                             b.raw(), (int64_t)n, k, b.stream()));
            } else {
                gpuErrChk(gputils_snullspace_project_batched((int)n, m_nullspace->raw(), n, nn, m_rank->raw(),
                                                             b.raw(), (int64_t)n, k, b.stream()));
            }
        } else {
            b.addAB(*m_projOp, b, 1, 0);
        }
    }
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
