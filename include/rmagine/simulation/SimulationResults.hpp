// Result attributes and bundles (standalone subset).  Same names and meaning as
// src/rmagine_core/include/rmagine/types/Bundle.hpp:72-132 (a struct inheriting from every
// attribute with a compile-time has<>() query) and
// src/rmagine_core/include/rmagine/simulation/SimulationResults.hpp:55-216 (attribute
// structs, IntAttrAll, resize_memory_bundle).
#pragma once

#include <cstdint>
#include <type_traits>

#include <rmagine/math/types.h>
#include <rmagine/types/Memory.hpp>

namespace rmagine {

template <typename... Tp>
struct Bundle : public Tp... {
  static constexpr std::size_t N = sizeof...(Tp);
  template <typename F>
  static constexpr bool has() {
    return std::disjunction<std::is_same<F, Tp>...>::value;
  }
};

template <typename MemT> struct Hits { Memory<uint8_t, MemT> hits; };
template <typename MemT> struct Ranges { Memory<float, MemT> ranges; };
template <typename MemT> struct Points { Memory<Point, MemT> points; };
template <typename MemT> struct Normals { Memory<Vector, MemT> normals; };
template <typename MemT> struct FaceIds { Memory<unsigned int, MemT> face_ids; };
template <typename MemT> using PrimitiveIds = FaceIds<MemT>;
template <typename MemT> struct GeomIds { Memory<unsigned int, MemT> geom_ids; };
template <typename MemT> using GeometryIds = GeomIds<MemT>;
template <typename MemT> struct ObjectIds { Memory<unsigned int, MemT> object_ids; };
template <typename MemT> using InstanceIds = ObjectIds<MemT>;

template <typename MemT>
using IntAttrAll = Bundle<Hits<MemT>, Ranges<MemT>, Points<MemT>, Normals<MemT>, FaceIds<MemT>, GeomIds<MemT>,
                          ObjectIds<MemT>>;
template <typename MemT>
using IntAttrAny = IntAttrAll<MemT>;

template <typename MemT, typename BundleT>
static void resize_memory_bundle(BundleT& res, unsigned int W, unsigned int H, unsigned int N = 1) {
  const size_t n = static_cast<size_t>(W) * H * N;
  if constexpr (BundleT::template has<Hits<MemT>>()) res.Hits<MemT>::hits.resize(n);
  if constexpr (BundleT::template has<Ranges<MemT>>()) res.Ranges<MemT>::ranges.resize(n);
  if constexpr (BundleT::template has<Points<MemT>>()) res.Points<MemT>::points.resize(n);
  if constexpr (BundleT::template has<Normals<MemT>>()) res.Normals<MemT>::normals.resize(n);
  if constexpr (BundleT::template has<FaceIds<MemT>>()) res.FaceIds<MemT>::face_ids.resize(n);
  if constexpr (BundleT::template has<GeomIds<MemT>>()) res.GeomIds<MemT>::geom_ids.resize(n);
  if constexpr (BundleT::template has<ObjectIds<MemT>>()) res.ObjectIds<MemT>::object_ids.resize(n);
}

}  // namespace rmagine
