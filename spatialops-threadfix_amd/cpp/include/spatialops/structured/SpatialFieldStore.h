// SpatialFieldStore.h -- temporary fields of the C++ front end
// (spatialops/structured/SpatialFieldStore.h:50-420).
//
// SpatialFieldStore::get / get_from_window hand out a SpatFldPtr to a new field
// whose storage comes from the engine's device pool (nebo_alloc: size buckets,
// freed blocks kept for reuse) or, on the CPU, from the heap.  The SpatFldPtr is
// reference-counted; when the last copy goes, the field's storage goes back to
// the pool, which is what the reference's store does with its Pool<T>.
#pragma once

#include <spatialops/structured/SpatialField.h>

#include <memory>

namespace SpatialOps {

template <typename FieldT>
class SpatFldPtr {
 public:
  SpatFldPtr() = default;
  // wraps `field` and takes ownership of it
  explicit SpatFldPtr(FieldT* field) : p_(field) {}
  SpatFldPtr(FieldT* field, bool builtFromStore) : p_(field), store_(builtFromStore) {}
  SpatFldPtr& operator=(FieldT* field)
  {
    p_.reset(field);
    store_ = false;
    return *this;
  }

  FieldT& operator*() { return *p_; }
  const FieldT& operator*() const { return *p_; }
  FieldT* operator->() { return p_.get(); }
  const FieldT* operator->() const { return p_.get(); }
  bool isnull() const { return !p_; }
  int count() const { return (int)p_.use_count(); }
  bool built_from_store() const { return store_; }
  // drop this reference (the storage goes back to the pool with the last one)
  void detach()
  {
    p_.reset();
    store_ = false;
  }

 private:
  std::shared_ptr<FieldT> p_;
  bool store_ = false;
};

class SpatialFieldStore {
 public:
  // a temporary with the given window, boundary info and ghosts on `deviceIndex`
  template <typename FieldT>
  static SpatFldPtr<FieldT> get_from_window(const MemoryWindow& w, const BoundaryCellInfo& bc,
                                           const GhostData& ghost, short deviceIndex = CPU_INDEX)
  {
    return SpatFldPtr<FieldT>(new FieldT(w, bc, ghost, nullptr, InternalStorage, deviceIndex), true);
  }

  // a temporary of type FieldT shaped after `f` (on f's active device by default)
  template <typename FieldT, typename ProtoT>
  static SpatFldPtr<FieldT> get(const ProtoT& f, short deviceIndex = -9999)
  {
    if (deviceIndex == -9999) deviceIndex = f.active_device_index();
    if (std::is_same<FieldT, SingleValueField>::value)
      return get_from_window<FieldT>(MemoryWindow(IntVec(1, 1, 1)), BoundaryCellInfo::build<FieldT>(),
                                     GhostData(0), deviceIndex);
    return get_from_window<FieldT>(create_new_memory_window<FieldT, ProtoT>(f),
                                   create_new_boundary_cell_info<FieldT, ProtoT>(f), f.get_ghost_data(),
                                   deviceIndex);
  }
};

}  // namespace SpatialOps
