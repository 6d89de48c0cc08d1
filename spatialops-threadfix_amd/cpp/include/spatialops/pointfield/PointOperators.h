// PointOperators.h -- PointToField and FieldToPoint of the C++ front end
// (spatialops/pointfield/PointOperators.h:28-160) on the engine's index scatter /
// gather (nebo_point_to_field / nebo_field_to_point).  The index list is
// uploaded once per device; where an index repeats, PointToField keeps the last
// point, as the reference's in-order loop does.
#pragma once

#include <spatialops/pointfield/PointFieldTypes.h>

#include <map>
#include <memory>
#include <vector>

namespace SpatialOps {
namespace Point {

namespace detail {

class PointMaps {
 public:
  explicit PointMaps(std::vector<size_t> idx) : idx_(std::move(idx)) {}
  template <typename FieldT>
  static std::vector<size_t> nonzero(const FieldT& locs)   // PointOperators.h:95-104
  {
    std::vector<size_t> out;
    size_t i = 0;
    for (typename FieldT::const_iterator it = locs.begin(); it != locs.end(); ++it, ++i)
      if (*it != 0.0) out.push_back(i);
    return out;
  }
  const nebo_point_map_t* on(short dev) const
  {
    auto it = maps_->find(dev);
    if (it != maps_->end()) return it->second.get();
    std::vector<int64_t> v(idx_.begin(), idx_.end());
    nebo_point_map_t* m = nullptr;
    nebo_check(nebo_point_map_create(v.data(), (int64_t)v.size(), dev, &m));
    (*maps_)[dev] = std::shared_ptr<nebo_point_map_t>(m, [](nebo_point_map_t* p) { nebo_point_map_destroy(p); });
    return m;
  }
  size_t size() const { return idx_.size(); }

 private:
  std::vector<size_t> idx_;
  std::shared_ptr<std::map<short, std::shared_ptr<nebo_point_map_t>>> maps_ =
      std::make_shared<std::map<short, std::shared_ptr<nebo_point_map_t>>>();
};

}  // namespace detail

template <typename FieldT>
class PointToField {
 public:
  typedef PointField SrcFieldType;
  typedef FieldT DestFieldType;
  explicit PointToField(const FieldT& fieldLocations) : m_(detail::PointMaps::nonzero(fieldLocations)) {}
  explicit PointToField(const std::vector<size_t>& flatIndices) : m_(flatIndices) {}
  void apply_to_field(const SrcFieldType& src, DestFieldType& dest) const
  {
    if (src.window_with_ghost().glob_npts() < m_.size())
      throw std::runtime_error("Nebo error in Nebo Point Field:\npoint field shorter than the index list");
    nebo_field_t d = dest.descriptor();
    nebo_check(nebo_point_to_field(m_.on((short)d.device), &d, src.field_values((short)d.device),
                                   dest.get_stream()));
    dest.mark_written();
  }

 private:
  detail::PointMaps m_;
};

template <typename FieldT>
class FieldToPoint {
 public:
  typedef FieldT SrcFieldType;
  typedef PointField DestFieldType;
  explicit FieldToPoint(const FieldT& fieldLocations) : m_(detail::PointMaps::nonzero(fieldLocations)) {}
  explicit FieldToPoint(const std::vector<size_t>& flatIndices) : m_(flatIndices) {}
  void apply_to_field(const SrcFieldType& src, DestFieldType& dest) const
  {
    if (dest.window_with_ghost().glob_npts() < m_.size())
      throw std::runtime_error("Nebo error in Nebo Point Field:\npoint field shorter than the index list");
    nebo_field_t s = src.descriptor();
    nebo_check(nebo_field_to_point(m_.on((short)s.device), dest.field_values((short)s.device), &s,
                                   dest.get_stream()));
    dest.mark_written();
  }

 private:
  detail::PointMaps m_;
};

}  // namespace Point
}  // namespace SpatialOps
