// Device data abstraction: the reference's DataContainer / DeviceContext interfaces
// (common/include/semblance/data/*.hpp) and their HIP implementations.
//
// DataContainer is a float-only device array whose size is counted in elements;
// copyFrom / pasteTo throw std::invalid_argument on a size mismatch and reset() zeroes it
// (CudaDataContainer.cu:31-63).  DeviceContext::activate() makes the device current; the
// HIP context also owns the stream all of its work is queued on (default priority).
// Non-float device storage the backend needs (trace samples in F16/BF16, index lists,
// masks) lives in DeviceBuffer.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

// Throws std::runtime_error for a failing HIP call (the reference maps CUDA errors the
// same way, CudaUtils.cu:8-15, so main() exits with -ENODEV).
void hipAssert(hipError_t code, const char *file, int line);
#define HIP_ASSERT(call) hipAssert((call), __FILE__, __LINE__)
// Same for a status of the C ABI (include/hss.h).
void hssAssert(int status, const char *file, int line);
#define HSS_ASSERT(call) hssAssert((call), __FILE__, __LINE__)

class DeviceContext {
  public:
    explicit DeviceContext(unsigned int devId) : deviceId(devId) {}
    virtual ~DeviceContext() = default;
    virtual void activate() const = 0;
    unsigned int getDeviceId() const { return deviceId; }

  protected:
    unsigned int deviceId;
};

class DeviceContextBuilder {
  public:
    virtual ~DeviceContextBuilder() = default;
    virtual DeviceContext *build(unsigned int devId) = 0;
};

class DataContainer {
  public:
    DataContainer(unsigned int count, std::shared_ptr<DeviceContext> context)
        : deviceContext(std::move(context)), elementCount(count) {}
    virtual ~DataContainer() = default;

    unsigned int getElementCount() const { return elementCount; }
    void reallocate(unsigned int newElementCount) {
        deallocate();
        elementCount = newElementCount;
        allocate();
    }

    virtual void allocate() = 0;
    virtual void deallocate() = 0;
    virtual void copyFrom(const std::vector<float> &source) = 0;
    virtual void copyFromWithOffset(const std::vector<float> &source, unsigned int offset) = 0;
    virtual void pasteTo(std::vector<float> &target) = 0;
    virtual void reset() = 0;

  protected:
    std::shared_ptr<DeviceContext> deviceContext;
    unsigned int elementCount;
};

class DataContainerBuilder {
  public:
    virtual ~DataContainerBuilder() = default;
    DataContainer *build(std::shared_ptr<DeviceContext> context) { return build(0, std::move(context)); }
    virtual DataContainer *build(unsigned int count, std::shared_ptr<DeviceContext> context) = 0;
};

// ---------------------------------------------------------------------------------
class HipDeviceContext : public DeviceContext {
  public:
    explicit HipDeviceContext(unsigned int devId);
    ~HipDeviceContext() override;
    void activate() const override;
    hipStream_t getStream() const { return stream; }
    void synchronize() const;

  private:
    hipStream_t stream = nullptr;
};

class HipDeviceContextBuilder : public DeviceContextBuilder {
  public:
    static DeviceContextBuilder *getInstance();
    DeviceContext *build(unsigned int devId) override { return new HipDeviceContext(devId); }
};

class HipDataContainer : public DataContainer {
  public:
    HipDataContainer(unsigned int count, std::shared_ptr<DeviceContext> context);
    ~HipDataContainer() override;
    float *getAddress() const { return address; }
    void allocate() override;
    void deallocate() override;
    void copyFrom(const std::vector<float> &source) override;
    void copyFromWithOffset(const std::vector<float> &source, unsigned int offset) override;
    void pasteTo(std::vector<float> &target) override;
    void reset() override;

  private:
    hipStream_t stream() const;
    float *address = nullptr;
};

class HipDataContainerBuilder : public DataContainerBuilder {
  public:
    static DataContainerBuilder *getInstance();
    using DataContainerBuilder::build;
    DataContainer *build(unsigned int count, std::shared_ptr<DeviceContext> context) override {
        return new HipDataContainer(count, std::move(context));
    }
};

#define HIP_DEV_PTR(container) (static_cast<HipDataContainer *>((container).get())->getAddress())

// Untyped device allocation owned by the backend.
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { resize(bytes); }
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void resize(size_t bytes);  // contents are not preserved
    void upload(const void *host, size_t bytes, hipStream_t stream);
    void download(void *host, size_t bytes, hipStream_t stream) const;
    void *get() const { return ptr; }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
    size_t size() const { return bytes_; }

  private:
    void *ptr = nullptr;
    size_t bytes_ = 0;
};
