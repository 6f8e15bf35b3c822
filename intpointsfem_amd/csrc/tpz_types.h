// tpz_types.h -- minimal value types standing in for the NeoPZ containers the
// reference's operator API is written against (TPZVec, TPZFMatrix,
// TPZSYsmpMatrix) plus a move-only device vector replacing TPZVecGPU
// (TPZVecGPU.h:9-111, whose implicit copy double-frees, TPZVecGPU.h:24-27).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

using REAL = double;
using STATE = double;

// Allocator whose value-less construct() default-initialises: resize(n) of a
// vector of doubles leaves the new elements unwritten, so a large buffer is
// first touched by the (parallel) code that fills it, not by a serial zero
// fill.
template <class T, class A = std::allocator<T>>
struct DefaultInitAllocator : A {
    template <class U>
    struct rebind {
        using other = DefaultInitAllocator<U, typename std::allocator_traits<A>::template rebind_alloc<U>>;
    };
    using A::A;
    template <class U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... Args>
    void construct(U* p, Args&&... args) {
        std::allocator_traits<A>::construct(static_cast<A&>(*this), p, std::forward<Args>(args)...);
    }
};

template <class T>
class TPZVec {
public:
    TPZVec() = default;
    explicit TPZVec(int64_t n, const T& v = T()) : fData(n, v) {}
    int64_t size() const { return (int64_t)fData.size(); }
    int64_t NElements() const { return size(); }
    // resize: new elements default-initialised (unwritten for arithmetic T);
    // Resize: new elements set to v
    void resize(int64_t n) { fData.resize(n); }
    void Resize(int64_t n, const T& v = T()) { fData.resize(n, v); }
    void Fill(const T& v) { std::fill(fData.begin(), fData.end(), v); }
    T& operator[](int64_t i) { return fData[i]; }
    const T& operator[](int64_t i) const { return fData[i]; }
    T* begin() { return fData.data(); }
    T* end() { return fData.data() + fData.size(); }
    const T* begin() const { return fData.data(); }
    const T* end() const { return fData.data() + fData.size(); }
    T* data() { return fData.data(); }
    const T* data() const { return fData.data(); }
    void push_back(const T& v) { fData.push_back(v); }

private:
    std::vector<T, DefaultInitAllocator<T>> fData;
};

// Column-major dense matrix (NeoPZ's TPZFMatrix storage order).
template <class T>
class TPZFMatrix {
public:
    TPZFMatrix() = default;
    TPZFMatrix(int64_t r, int64_t c, const T& v = T()) : fRows(r), fCols(c), fData(r * c, v) {}
    int64_t Rows() const { return fRows; }
    int64_t Cols() const { return fCols; }
    void Resize(int64_t r, int64_t c) {
        std::vector<T> n(r * c, T());
        for (int64_t j = 0; j < std::min(c, fCols); j++)
            for (int64_t i = 0; i < std::min(r, fRows); i++) n[j * r + i] = fData[j * fRows + i];
        fRows = r;
        fCols = c;
        fData.swap(n);
    }
    void Redim(int64_t r, int64_t c) {
        fRows = r;
        fCols = c;
        fData.assign(r * c, T());
    }
    void Zero() { std::fill(fData.begin(), fData.end(), T()); }
    T& operator()(int64_t i, int64_t j = 0) { return fData[j * fRows + i]; }
    const T& operator()(int64_t i, int64_t j = 0) const { return fData[j * fRows + i]; }
    TPZFMatrix& operator+=(const TPZFMatrix& o) {
        if (o.fRows != fRows || o.fCols != fCols) throw std::invalid_argument("TPZFMatrix += size mismatch");
        for (size_t i = 0; i < fData.size(); i++) fData[i] = fData[i] + o.fData[i];
        return *this;
    }
    T* data() { return fData.data(); }
    const T* data() const { return fData.data(); }

private:
    int64_t fRows = 0, fCols = 0;
    std::vector<T> fData;
};

// Base of NeoPZ's matrices: what TPZStructMatrix::Create() returns and
// Assemble(mat, rhs, gui) takes.
template <class T>
class TPZMatrix {
public:
    virtual ~TPZMatrix() = default;
    virtual int64_t Rows() const = 0;
    virtual int64_t Cols() const { return Rows(); }
};

// Symmetric sparse matrix, upper triangle in CSR (NeoPZ TPZSYsmpMatrix:
// IA row pointers, JA column indexes, A values).
template <class T>
class TPZSYsmpMatrix : public TPZMatrix<T> {
public:
    TPZVec<int64_t>& IA() { return fIA; }
    TPZVec<int64_t>& JA() { return fJA; }
    TPZVec<T>& A() { return fA; }
    int64_t Rows() const override { return fIA.size() ? fIA.size() - 1 : 0; }
    T GetVal(int64_t r, int64_t c) const {
        if (r > c) std::swap(r, c);
        for (int64_t k = fIA[r]; k < fIA[r + 1]; k++)
            if (fJA[k] == c) return fA[k];
        return T();
    }
    void PutVal(int64_t r, int64_t c, const T& v) {
        if (r > c) std::swap(r, c);
        for (int64_t k = fIA[r]; k < fIA[r + 1]; k++)
            if (fJA[k] == c) {
                fA[k] = v;
                return;
            }
        throw std::out_of_range("TPZSYsmpMatrix::PutVal outside the pattern");
    }

private:
    TPZVec<int64_t> fIA, fJA;
    TPZVec<T> fA;
};

// NeoPZ's progress / cancel hook passed through the assembly calls; the
// integration-point path never consults it.
class TPZGuiInterface {
public:
    virtual ~TPZGuiInterface() = default;
};

// NeoPZ's reference-counted pointer (shared ownership).
template <class T>
class TPZAutoPointer {
public:
    TPZAutoPointer() = default;
    explicit TPZAutoPointer(T* p) : fRef(p) {}
    T* operator->() const { return fRef.get(); }
    T& operator*() const { return *fRef; }
    explicit operator bool() const { return (bool)fRef; }

private:
    std::shared_ptr<T> fRef;
};

// RAII device vector, move-only.
template <class T>
class TPZVecGPU {
public:
    TPZVecGPU() = default;
    explicit TPZVecGPU(int64_t n) { resize(n); }
    ~TPZVecGPU() { release(); }
    TPZVecGPU(const TPZVecGPU&) = delete;
    TPZVecGPU& operator=(const TPZVecGPU&) = delete;
    TPZVecGPU(TPZVecGPU&& o) noexcept : fData(o.fData), fSize(o.fSize) {
        o.fData = nullptr;
        o.fSize = 0;
    }
    TPZVecGPU& operator=(TPZVecGPU&& o) noexcept {
        if (this != &o) {
            release();
            fData = o.fData;
            fSize = o.fSize;
            o.fData = nullptr;
            o.fSize = 0;
        }
        return *this;
    }
    void resize(int64_t n) {
        if (n == fSize) return;
        release();
        if (n > 0) check(hipMalloc((void**)&fData, sizeof(T) * n), "hipMalloc");
        fSize = n;
    }
    void set(const T* host, int64_t n) {
        resize(n);
        if (n) check(hipMemcpy(fData, host, sizeof(T) * n, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void get(T* host, int64_t n) const {
        if (n > fSize) throw std::out_of_range("TPZVecGPU::get beyond size");
        if (n) check(hipMemcpy(host, fData, sizeof(T) * n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
    void Zero() {
        if (fSize) check(hipMemset(fData, 0, sizeof(T) * fSize), "hipMemset");
    }
    T* getData() { return fData; }
    const T* getData() const { return fData; }
    int64_t getSize() const { return fSize; }

private:
    static void check(hipError_t e, const char* what) {
        if (e != hipSuccess) throw std::runtime_error(std::string(what) + " failed: " + hipGetErrorString(e));
    }
    void release() {
        if (fData) (void)hipFree(fData);
        fData = nullptr;
        fSize = 0;
    }
    T* fData = nullptr;
    int64_t fSize = 0;
};
