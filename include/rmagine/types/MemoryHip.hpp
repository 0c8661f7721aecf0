// VRAM_HIP memory tag: device memory of librmh (the HIP counterpart of VRAM_CUDA,
// src/rmagine_core/include/rmagine/types/MemoryCuda.hpp:66-98 and
// src/rmagine_cuda/src/types/MemoryCuda.cu:13-89).  Copies are blocking, as in the reference.
#pragma once

#include <rmh.h>

#include <rmagine/types/Memory.hpp>
#include <rmagine/util/hip/HipException.hpp>

namespace rmagine {

struct VRAM_HIP {
  template <typename DataT>
  static DataT* alloc(size_t N) {
    void* p = nullptr;
    RM_RMH_CHECK(rmh_malloc(&p, sizeof(DataT) * N));
    return static_cast<DataT*>(p);
  }
  template <typename DataT>
  static DataT* realloc(DataT* mem, size_t Nold, size_t Nnew) {
    // contents are not preserved (MemoryCuda.tcc:15-22)
    free(mem, Nold);
    return alloc<DataT>(Nnew);
  }
  template <typename DataT>
  static void free(DataT* mem, size_t) {
    if (mem) (void)rmh_free(mem);
  }
};

template <typename DataT>
void copy(const MemoryView<DataT, RAM>& from, MemoryView<DataT, VRAM_HIP>& to) {
  RM_RMH_CHECK(rmh_memcpy(to.raw(), from.raw(), sizeof(DataT) * from.size(), 0));
}

template <typename DataT>
void copy(const MemoryView<DataT, VRAM_HIP>& from, MemoryView<DataT, RAM>& to) {
  RM_RMH_CHECK(rmh_memcpy(to.raw(), from.raw(), sizeof(DataT) * from.size(), 1));
}

template <typename DataT>
void copy(const MemoryView<DataT, VRAM_HIP>& from, MemoryView<DataT, VRAM_HIP>& to) {
  RM_RMH_CHECK(rmh_memcpy(to.raw(), from.raw(), sizeof(DataT) * from.size(), 2));
}

}  // namespace rmagine
