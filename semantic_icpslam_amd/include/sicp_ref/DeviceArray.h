// DeviceArray.h — the reference's device buffer types (Utility/DeviceArray.h:9-403)
// over HIP, for callers written against the reference.
//
// Same public surface as the reference: `data`, `size` / `step`, `cols`,
// `rows` fields; create / upload / download / clear / release / swap /
// copyTo; shallow copies that share the allocation; 2-D buffers pitched, row
// y at (char*)data + y * step. Ownership is a shared_ptr instead of a hand
// reference count. Errors go through SafeCall (SafeCall.h in this directory).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <memory>
#include <vector>

#include "SafeCall.h"
#include "sicp.h"

template <class T> struct DeviceArray {
    DeviceArray() = default;
    explicit DeviceArray(size_t size_) { create(size_); }
    explicit DeviceArray(const std::vector<T>& vec) {
        create(vec.size());
        upload(vec);
    }

    void create(size_t size_) {
        release();
        void* p = nullptr;
        SafeCall(hipMalloc(&p, sizeof(T) * (size_ ? size_ : 1)));
        own_ = std::shared_ptr<void>(p, [](void* q) { (void)hipFree(q); });
        data = p;
        size = size_;
    }
    // A view of device memory owned elsewhere (no free on release).
    void wrap(void* device_ptr, size_t size_) {
        release();
        data = device_ptr;
        size = size_;
    }
    void upload(const void* data_) { upload(data_, size); }
    void upload(const std::vector<T>& vec) { upload(vec.data(), vec.size()); }
    // Larger than the buffer: ignored, as in the reference (DeviceArray.h:201-204).
    void upload(const void* data_, size_t size_) {
        if (size_ > size) return;
        SafeCall(hipMemcpy(data, data_, sizeof(T) * size_, hipMemcpyHostToDevice));
    }
    void download(void* data_) const { download(data_, size); }
    void download(std::vector<T>& vec) const {
        vec.resize(size);
        download(vec.data(), vec.size());
    }
    void download(void* data_, size_t size_) const {
        SafeCall(hipMemcpy(data_, data, sizeof(T) * size_, hipMemcpyDeviceToHost));
    }
    void clear() { SafeCall(hipMemset(data, 0, sizeof(T) * size)); }
    void release() {
        own_.reset();
        data = nullptr;
        size = 0;
    }
    void copyTo(DeviceArray<T>& other) const {
        if (!data) {
            other.release();
            return;
        }
        other.create(size);
        SafeCall(hipMemcpy(other.data, data, sizeof(T) * size, hipMemcpyDeviceToDevice));
    }
    operator T*() const { return (T*)data; }

    void* data = nullptr;
    size_t size = 0;

  private:
    std::shared_ptr<void> own_;
};

template <class T> struct DeviceArray2D {
    DeviceArray2D() = default;
    DeviceArray2D(int cols_, int rows_) { create(cols_, rows_); }

    void create(int cols_, int rows_) {
        if (cols_ <= 0 || rows_ <= 0) return;
        release();
        void* p = nullptr;
        size_t pitch = 0;
        SafeCall(hipMallocPitch(&p, &pitch, sizeof(T) * cols_, rows_));
        own_ = std::shared_ptr<void>(p, [](void* q) { (void)hipFree(q); });
        data = p;
        step = pitch;
        cols = cols_;
        rows = rows_;
    }
    void upload(const void* data_) { upload(data_, sizeof(T) * cols, cols, rows); }
    void upload(const void* data_, size_t step_) { upload(data_, step_, cols, rows); }
    void upload(const void* data_, size_t step_, int cols_, int rows_) {
        if (!data) create(cols_, rows_);
        SafeCall(hipMemcpy2D(data, step, data_, step_, sizeof(T) * cols_, rows_,
                             hipMemcpyHostToDevice));
    }
    void download(void* data_, size_t step_) const {
        if (!data) return;
        SafeCall(hipMemcpy2D(data_, step_, data, step, sizeof(T) * cols, rows,
                             hipMemcpyDeviceToHost));
    }
    void clear() { SafeCall(hipMemset2D(data, step, 0, sizeof(T) * cols, rows)); }
    void release() {
        own_.reset();
        data = nullptr;
        step = 0;
        cols = rows = 0;
    }
    void swap(DeviceArray2D<T>& other) {
        std::swap(own_, other.own_);
        std::swap(data, other.data);
        std::swap(step, other.step);
        std::swap(cols, other.cols);
        std::swap(rows, other.rows);
    }
    void copyTo(DeviceArray2D<T>& other) const {
        if (!data) {
            other.release();
            return;
        }
        other.create(cols, rows);
        SafeCall(hipMemcpy2D(other.data, other.step, data, step, sizeof(T) * cols, rows,
                             hipMemcpyDeviceToDevice));
    }
    operator T*() const { return (T*)data; }
    T* ptr(int y = 0) const { return (T*)((char*)data + y * step); }

    // The C ABI's descriptor of this buffer.
    sicp_image image() const { return sicp_image{data, step, cols, rows}; }

    void* data = nullptr;
    size_t step = 0;
    int cols = 0, rows = 0;

  private:
    std::shared_ptr<void> own_;
};
