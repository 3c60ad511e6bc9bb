// Stand-in of the parts of linearAlgebra::MultiVector that the drop-in
// header uses (src/linearAlgebra/MultiVector.h:191-487: a copy constructor,
// the copy constructor with an initial value, data(), getNumberComponents(),
// localSize(), locallyOwnedSize()).  DEVICE storage is HIP device memory, as
// dft-efe's own device MultiVector; one process, so no ghost rows.
#pragma once
#include <hip/hip_runtime_api.h>
#include <utils/MemorySpaceType.h>
#include <utils/TypeConfig.h>

#include <algorithm>
#include <stdexcept>
#include <vector>
namespace dftefe
{
  namespace linearAlgebra
  {
    template <typename ValueType, utils::MemorySpace memorySpace>
    class MultiVector
    {
      static constexpr bool dev = memorySpace == utils::MemorySpace::DEVICE;

    public:
      MultiVector(size_type locallyOwnedSize, size_type numVectors, ValueType v = ValueType())
        : d_n(locallyOwnedSize), d_b(numVectors)
      {
        alloc();
        fill(v);
      }
      MultiVector(const MultiVector &u, ValueType initVal) : d_n(u.d_n), d_b(u.d_b)
      {
        alloc();
        fill(initVal);
      }
      MultiVector(const MultiVector &u) : d_n(u.d_n), d_b(u.d_b)
      {
        alloc();
        copy(u);
      }
      MultiVector &operator=(const MultiVector &u)
      {
        if (this != &u)
          {
            release();
            d_n = u.d_n;
            d_b = u.d_b;
            alloc();
            copy(u);
          }
        return *this;
      }
      ~MultiVector() { release(); }
      ValueType *data() { return d_p; }
      const ValueType *data() const { return d_p; }
      size_type getNumberComponents() const { return d_b; }
      size_type localSize() const { return d_n; }
      size_type locallyOwnedSize() const { return d_n; }

    private:
      size_t count() const { return (size_t)d_n * d_b; }
      void alloc()
      {
        if (dev)
          {
            if (hipMalloc(reinterpret_cast<void **>(&d_p), count() * sizeof(ValueType)) !=
                hipSuccess)
              throw std::runtime_error("hipMalloc failed");
          }
        else
          d_p = new ValueType[count()];
      }
      void release()
      {
        if (!d_p)
          return;
        if (dev)
          (void)hipFree(d_p);
        else
          delete[] d_p;
        d_p = nullptr;
      }
      void fill(ValueType v)
      {
        std::vector<ValueType> h(count(), v);
        if (dev)
          (void)hipMemcpy(d_p, h.data(), h.size() * sizeof(ValueType), hipMemcpyHostToDevice);
        else
          std::copy(h.begin(), h.end(), d_p);
      }
      void copy(const MultiVector &u)
      {
        if (dev)
          (void)hipMemcpy(d_p, u.d_p, count() * sizeof(ValueType), hipMemcpyDeviceToDevice);
        else
          std::copy(u.d_p, u.d_p + count(), d_p);
      }
      size_type  d_n, d_b;
      ValueType *d_p = nullptr;
    };
  } // namespace linearAlgebra
} // namespace dftefe
