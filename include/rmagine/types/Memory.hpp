// Standalone subset of Rmagine's memory types for builds without rmagine_core.
//
// Same shape and semantics as src/rmagine_core/include/rmagine/types/Memory.hpp:59-289 and
// Memory.tcc:15-226: MemoryView is a non-owning (pointer, size) pair tagged with its
// memory space; Memory owns its buffer, deep-copies on copy, steals on move, and
// assigning a view of another memory space resizes and copies across spaces.  Inside an
// Rmagine tree the reference's own Memory.hpp takes this file's place; MemoryHip.hpp only
// adds the VRAM_HIP tag and its copy overloads.
#pragma once

#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <utility>

namespace rmagine {

struct RAM {
  template <typename DataT>
  static DataT* alloc(size_t N) { return N ? new DataT[N] : nullptr; }
  template <typename DataT>
  static DataT* realloc(DataT* mem, size_t Nold, size_t Nnew) {
    DataT* n = alloc<DataT>(Nnew);
    if (mem && n) std::memcpy(n, mem, sizeof(DataT) * (Nold < Nnew ? Nold : Nnew));
    free(mem, Nold);
    return n;
  }
  template <typename DataT>
  static void free(DataT* mem, size_t) { delete[] mem; }
};

template <typename DataT, typename MemT = RAM>
class MemoryView {
 public:
  using DataType = DataT;
  using MemType = MemT;

  MemoryView() = delete;
  MemoryView(DataT* mem, size_t N) : m_mem(mem), m_size(N) {}
  // copying a view copies the (pointer, size) pair; assignment copies the elements
  MemoryView(const MemoryView&) = default;

  static MemoryView<DataT, MemT> Empty() { return MemoryView<DataT, MemT>(nullptr, 0); }

  bool empty() const { return m_mem == nullptr; }
  DataT* raw() { return m_mem; }
  const DataT* raw() const { return m_mem; }
  size_t size() const { return m_size; }

  DataT& operator[](size_t i) { return m_mem[i]; }
  const DataT& operator[](size_t i) const { return m_mem[i]; }
  DataT& at(size_t i) { return m_mem[i]; }
  const DataT& at(size_t i) const { return m_mem[i]; }

  MemoryView<DataT, MemT> slice(size_t a, size_t b) { return MemoryView<DataT, MemT>(m_mem + a, b - a); }
  const MemoryView<DataT, MemT> slice(size_t a, size_t b) const {
    return MemoryView<DataT, MemT>(m_mem + a, b - a);
  }
  MemoryView<DataT, MemT> operator()(size_t a, size_t b) { return slice(a, b); }
  const MemoryView<DataT, MemT> operator()(size_t a, size_t b) const { return slice(a, b); }

  MemoryView<DataT, MemT>& operator=(const MemoryView<DataT, MemT>& o);
  template <typename MemT2>
  MemoryView<DataT, MemT>& operator=(const MemoryView<DataT, MemT2>& o);

 protected:
  DataT* m_mem = nullptr;
  size_t m_size = 0;
};

template <typename DataT, typename MemT = RAM>
using MemView = MemoryView<DataT, MemT>;

template <typename T>
MemoryView<T, RAM> make_view(T& data) { return MemoryView<T, RAM>(&data, 1); }

// copy between memory spaces: RAM<->RAM here, VRAM_HIP overloads in MemoryHip.hpp
template <typename DataT>
void copy(const MemoryView<DataT, RAM>& from, MemoryView<DataT, RAM>& to) {
  if (from.size()) std::memcpy(to.raw(), from.raw(), sizeof(DataT) * from.size());
}

template <typename DataT, typename MemT>
MemoryView<DataT, MemT>& MemoryView<DataT, MemT>::operator=(const MemoryView<DataT, MemT>& o) {
  if (o.size() != size()) throw std::runtime_error("MemoryView assignment of different sizes");
  copy(o, *this);
  return *this;
}

template <typename DataT, typename MemT>
template <typename MemT2>
MemoryView<DataT, MemT>& MemoryView<DataT, MemT>::operator=(const MemoryView<DataT, MemT2>& o) {
  if (o.size() != size()) throw std::runtime_error("MemoryView assignment of different sizes");
  copy(o, *this);
  return *this;
}

template <typename DataT, typename MemT = RAM>
class Memory : public MemoryView<DataT, MemT> {
 public:
  using Base = MemoryView<DataT, MemT>;

  Memory() : Base(nullptr, 0) {}
  explicit Memory(size_t N) : Base(nullptr, 0) { resize(N); }
  Memory(const Memory<DataT, MemT>& o) : Base(nullptr, 0) { *this = static_cast<const Base&>(o); }
  Memory(const MemoryView<DataT, MemT>& o) : Base(nullptr, 0) { *this = o; }
  template <typename MemT2>
  Memory(const MemoryView<DataT, MemT2>& o) : Base(nullptr, 0) { *this = o; }
  Memory(Memory<DataT, MemT>&& o) noexcept : Base(o.m_mem, o.m_size) {
    o.m_mem = nullptr;
    o.m_size = 0;
  }
  ~Memory() {
    if (m_mem) MemT::template free<DataT>(m_mem, m_size);
  }

  // Memory::resize (Memory.tcc:126-137): the memory space's realloc keeps the first
  // min(old, new) elements in RAM and drops the contents of device memory, as
  // VRAM_CUDA::realloc does (MemoryCuda.tcc:15-22)
  void resize(size_t N) {
    if (N == m_size) return;
    if (m_mem) m_mem = MemT::template realloc<DataT>(m_mem, m_size, N);
    else m_mem = MemT::template alloc<DataT>(N);
    m_size = N;
  }

  Memory<DataT, MemT>& operator=(const MemoryView<DataT, MemT>& o) {
    if (o.raw() == m_mem) return *this;
    resize(o.size());
    copy(o, static_cast<Base&>(*this));
    return *this;
  }
  Memory<DataT, MemT>& operator=(const Memory<DataT, MemT>& o) { return operator=(static_cast<const Base&>(o)); }
  Memory<DataT, MemT>& operator=(Memory<DataT, MemT>&& o) noexcept {
    std::swap(m_mem, o.m_mem);
    std::swap(m_size, o.m_size);
    return *this;
  }
  template <typename MemT2>
  Memory<DataT, MemT>& operator=(const MemoryView<DataT, MemT2>& o) {
    resize(o.size());
    copy(o, static_cast<Base&>(*this));
    return *this;
  }

 protected:
  using Base::m_mem;
  using Base::m_size;
};

template <typename DataT, typename MemT = RAM>
using Mem = Memory<DataT, MemT>;

}  // namespace rmagine
