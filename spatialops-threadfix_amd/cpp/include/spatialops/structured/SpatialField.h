// SpatialField.h -- fields, memory windows, ghost data and boundary-cell info of
// the C++ front end.  Same names and meaning as the reference's
// spatialops/structured/{SpatialField,MemoryWindow,GhostData,BoundaryCellInfo,
// FVStaggeredLocationTypes,FVStaggeredFieldTypes}.h; storage lives on the host
// and/or on GPUs through the engine's C-ABI (include/nebo_hip.h).
#pragma once

#include <nebo_hip.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#define CPU_INDEX (-1)
#define GPU_INDEX 0
#define IS_CPU_INDEX(i) ((i) == CPU_INDEX)
#define IS_GPU_INDEX(i) ((i) >= 0)
#define GHOST_MAX NEBO_GHOST_MAX

namespace SpatialOps {

// Rethrow a C-ABI failure as the reference's std::runtime_error("Nebo error in ...").
inline void nebo_check(int rc)
{
  if (rc != NEBO_OK) throw std::runtime_error(nebo_last_error());
}

// ---------------------------------------------------------------------------
// IntVec (spatialops/structured/IntVec.h)
// ---------------------------------------------------------------------------
struct IntVec {
  int v[3];
  IntVec(int x = 0, int y = 0, int z = 0) : v{x, y, z} {}
  int& operator[](int i) { return v[i]; }
  int operator[](int i) const { return v[i]; }
  IntVec operator+(const IntVec& o) const { return IntVec(v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]); }
  IntVec operator-(const IntVec& o) const { return IntVec(v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]); }
  IntVec operator*(const IntVec& o) const { return IntVec(v[0] * o.v[0], v[1] * o.v[1], v[2] * o.v[2]); }
  IntVec operator+(int s) const { return IntVec(v[0] + s, v[1] + s, v[2] + s); }
  bool operator==(const IntVec& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
  bool operator!=(const IntVec& o) const { return !(*this == o); }
};
inline std::ostream& operator<<(std::ostream& os, const IntVec& v)
{
  return os << "[" << v[0] << "," << v[1] << "," << v[2] << "]";
}

struct XDIR { enum { value = 0 }; };
struct YDIR { enum { value = 1 }; };
struct ZDIR { enum { value = 2 }; };
struct NODIR { enum { value = -1 }; };

// ---------------------------------------------------------------------------
// GhostData (GhostData.h:54-257): counts >= GHOST_MAX on all six faces = infinite
// ---------------------------------------------------------------------------
class GhostData {
  IntVec minus_, plus_;

 public:
  GhostData(int n = 0) : minus_(n, n, n), plus_(n, n, n) {}
  GhostData(int nx, int px, int ny, int py, int nz, int pz) : minus_(nx, ny, nz), plus_(px, py, pz) {}
  GhostData(const IntVec& m, const IntVec& p) : minus_(m), plus_(p) {}
  IntVec get_minus() const { return minus_; }
  IntVec get_plus() const { return plus_; }
  int get_minus(int i) const { return minus_[i]; }
  int get_plus(int i) const { return plus_[i]; }
  void set_minus(const IntVec& m) { minus_ = m; }
  void set_plus(const IntVec& p) { plus_ = p; }
  bool operator==(const GhostData& o) const { return minus_ == o.minus_ && plus_ == o.plus_; }
  GhostData limit_by_extent(const IntVec& e) const
  {
    return GhostData(e[0] <= 1 ? 0 : minus_[0], e[0] <= 1 ? 0 : plus_[0], e[1] <= 1 ? 0 : minus_[1],
                     e[1] <= 1 ? 0 : plus_[1], e[2] <= 1 ? 0 : minus_[2], e[2] <= 1 ? 0 : plus_[2]);
  }
};

// ---------------------------------------------------------------------------
// Locations (FVStaggeredLocationTypes.h:63-126): staggering offset and the
// extra cell a field gets on a (+) face that carries a boundary condition.
// ---------------------------------------------------------------------------
#define NEBO_LOCATION(NAME, DIR, OX, OY, OZ, EX, EY, EZ)                          \
  struct NAME {                                                                  \
    typedef DIR FaceDir;                                                         \
    static IntVec offset() { return IntVec(OX, OY, OZ); }                        \
    static IntVec bc_extra() { return IntVec(EX, EY, EZ); }                      \
    static constexpr int ox = OX, oy = OY, oz = OZ;                              \
  };
NEBO_LOCATION(SVol, NODIR, 0, 0, 0, 0, 0, 0)
NEBO_LOCATION(SSurfX, XDIR, -1, 0, 0, 1, 0, 0)
NEBO_LOCATION(SSurfY, YDIR, 0, -1, 0, 0, 1, 0)
NEBO_LOCATION(SSurfZ, ZDIR, 0, 0, -1, 0, 0, 1)
NEBO_LOCATION(XVol, NODIR, -1, 0, 0, 1, 0, 0)
NEBO_LOCATION(XSurfX, XDIR, 0, 0, 0, 0, 0, 0)
NEBO_LOCATION(XSurfY, YDIR, -1, -1, 0, 0, 1, 0)
NEBO_LOCATION(XSurfZ, ZDIR, -1, 0, -1, 0, 0, 1)
NEBO_LOCATION(YVol, NODIR, 0, -1, 0, 0, 1, 0)
NEBO_LOCATION(YSurfX, XDIR, -1, -1, 0, 1, 0, 0)
NEBO_LOCATION(YSurfY, YDIR, 0, 0, 0, 0, 0, 0)
NEBO_LOCATION(YSurfZ, ZDIR, 0, -1, -1, 0, 0, 1)
NEBO_LOCATION(ZVol, NODIR, 0, 0, -1, 0, 0, 1)
NEBO_LOCATION(ZSurfX, XDIR, -1, 0, -1, 1, 0, 0)
NEBO_LOCATION(ZSurfY, YDIR, 0, -1, -1, 0, 1, 0)
NEBO_LOCATION(ZSurfZ, ZDIR, 0, 0, 0, 0, 0, 0)
NEBO_LOCATION(SingleValue, NODIR, 0, 0, 0, 0, 0, 0)
#undef NEBO_LOCATION

// ---------------------------------------------------------------------------
// BoundaryCellInfo (BoundaryCellInfo.h:45-158)
// ---------------------------------------------------------------------------
class BoundaryCellInfo {
  IntVec isbc_, nExtra_;
  BoundaryCellInfo(bool bx, bool by, bool bz, const IntVec& extra)
      : isbc_(bx, by, bz), nExtra_(IntVec(bx, by, bz) * extra) {}

 public:
  template <typename FieldT>
  static BoundaryCellInfo build(bool bx, bool by, bool bz)
  {
    return BoundaryCellInfo(bx, by, bz, FieldT::Location::bc_extra());
  }
  template <typename FieldT>
  static BoundaryCellInfo build(const IntVec& bc)
  {
    return build<FieldT>(bc[0] != 0, bc[1] != 0, bc[2] != 0);
  }
  template <typename FieldT>
  static BoundaryCellInfo build() { return build<FieldT>(false, false, false); }
  bool has_bc(int d) const { return isbc_[d] != 0; }
  IntVec has_bc() const { return isbc_; }
  int num_extra(int d) const { return nExtra_[d]; }
  IntVec num_extra() const { return nExtra_; }
  int has_extra(int d) const { return has_bc(d) ? nExtra_[d] : 0; }
  IntVec has_extra() const { return IntVec(has_extra(0), has_extra(1), has_extra(2)); }
  BoundaryCellInfo limit_by_extent(const IntVec& e) const
  {
    BoundaryCellInfo b(*this);
    for (int d = 0; d < 3; ++d)
      if (e[d] == 1) b.nExtra_[d] = 0;
    return b;
  }
};

// ---------------------------------------------------------------------------
// MemoryWindow (MemoryWindow.h:63-321)
// ---------------------------------------------------------------------------
class MemoryWindow {
  IntVec glob_, offset_, extent_;

 public:
  explicit MemoryWindow(const IntVec& glob) : glob_(glob), offset_(0, 0, 0), extent_(glob) {}
  MemoryWindow(const IntVec& glob, const IntVec& offset, const IntVec& extent)
      : glob_(glob), offset_(offset), extent_(extent) {}
  const IntVec& glob_dim() const { return glob_; }
  const IntVec& offset() const { return offset_; }
  const IntVec& extent() const { return extent_; }
  size_t glob_dim(int i) const { return (size_t)glob_[i]; }
  size_t offset(int i) const { return (size_t)offset_[i]; }
  size_t extent(int i) const { return (size_t)extent_[i]; }
  size_t glob_npts() const { return (size_t)glob_[0] * glob_[1] * glob_[2]; }
  size_t local_npts() const { return (size_t)extent_[0] * extent_[1] * extent_[2]; }
  int flat_index(IntVec loc) const
  {
    for (int d = 0; d < 3; ++d) loc[d] = glob_[d] > 1 ? loc[d] + offset_[d] : 0;
    return loc[0] + glob_[0] * (loc[1] + glob_[1] * loc[2]);
  }
  bool operator==(const MemoryWindow& o) const
  {
    return glob_ == o.glob_ && offset_ == o.offset_ && extent_ == o.extent_;
  }
  bool operator!=(const MemoryWindow& o) const { return !(*this == o); }
};

inline int get_dim_with_ghost(int n, int gm, int gp, int bc) { return n > 1 ? n + gm + gp + bc : 1; }

// Host iterator over a window, x fastest then y then z (FieldIterator,
// MemoryWindow.h:330-560): the i-th point of the window's extent.
template <typename T>
class WindowIterator {
 public:
  WindowIterator(T* values, const MemoryWindow& w, size_t i = 0) : v_(values), w_(w), i_(i) {}
  T& operator*() const
  {
    const size_t ex = w_.extent(0), ey = w_.extent(1);
    const size_t x = i_ % ex, y = (i_ / ex) % ey, z = i_ / (ex * ey);
    return v_[(w_.offset(0) + x) + w_.glob_dim(0) * ((w_.offset(1) + y) + w_.glob_dim(1) * (w_.offset(2) + z))];
  }
  WindowIterator& operator++()
  {
    ++i_;
    return *this;
  }
  WindowIterator operator++(int)
  {
    WindowIterator t(*this);
    ++i_;
    return t;
  }
  WindowIterator operator+(size_t n) const { return WindowIterator(v_, w_, i_ + n); }
  bool operator==(const WindowIterator& o) const { return v_ == o.v_ && i_ == o.i_; }
  bool operator!=(const WindowIterator& o) const { return !(*this == o); }

 private:
  T* v_;
  MemoryWindow w_;
  size_t i_;
};

inline MemoryWindow get_window_with_ghost(const IntVec& n, const GhostData& g, const BoundaryCellInfo& bc)
{
  return MemoryWindow(IntVec(get_dim_with_ghost(n[0], g.get_minus(0), g.get_plus(0), bc.has_extra(0)),
                             get_dim_with_ghost(n[1], g.get_minus(1), g.get_plus(1), bc.has_extra(1)),
                             get_dim_with_ghost(n[2], g.get_minus(2), g.get_plus(2), bc.has_extra(2))));
}

enum StorageMode { InternalStorage, ExternalStorage };

// ---------------------------------------------------------------------------
// Field storage with one copy per device and validity flags (FieldInfo.h:91-795)
// ---------------------------------------------------------------------------
template <typename T>
class FieldStorage {
 public:
  struct Copy {
    T* ptr = nullptr;
    bool valid = false;
    bool owned = false;
    bool pinned = false;   // host copy in page-locked memory
  };

  FieldStorage(size_t n, T* external, StorageMode mode, short dev) : n_(n), active_(dev)
  {
    Copy& c = copies_[dev];
    if (mode == ExternalStorage && external) {
      c.ptr = external;
    } else {
      allocate(dev, c);
    }
    c.valid = true;
  }
  ~FieldStorage()
  {
    for (auto& kv : copies_)
      if (kv.second.owned) release(kv.first, kv.second);
  }
  FieldStorage(const FieldStorage&) = delete;
  FieldStorage& operator=(const FieldStorage&) = delete;

  size_t npts() const { return n_; }
  short active() const { return active_; }
  bool is_valid(short d) const
  {
    auto it = copies_.find(d);
    return it != copies_.end() && it->second.valid;
  }
  bool is_available(short d) const { return copies_.count(d) != 0; }

  // Make the copy on d valid.  With a stream the copy is queued on it and the
  // call returns at once (the *_async forms of FieldInfo.h:509-678): work queued
  // later on that stream sees the data, the host waits with a stream sync.
  void add_device(short d, void* stream = nullptr)
  {
    std::lock_guard<std::mutex> lk(mu_);
    Copy& c = copies_[d];
    if (!c.ptr) allocate(d, c);
    if (!c.valid) {
      copy(d, c.ptr, active_, copies_[active_].ptr, stream);
      c.valid = true;
    }
  }
  void validate_device(short d, void* stream = nullptr) { add_device(d, stream); }
  void set_device_as_active(short d, void* stream = nullptr)
  {
    add_device(d, stream);
    active_ = d;
  }
  // write access: must be the active device; other copies become stale
  T* values(short d)
  {
    if (d != active_)
      throw std::runtime_error("Nebo error in FieldStorage:\nrequested write access to a device that is not active");
    for (auto& kv : copies_)
      if (kv.first != d) kv.second.valid = false;
    return copies_[d].ptr;
  }
  const T* const_values(short d) const
  {
    auto it = copies_.find(d);
    if (it == copies_.end() || !it->second.valid)
      throw std::runtime_error("Nebo error in FieldStorage:\nrequested read access to a device without a valid copy");
    return it->second.ptr;
  }
  T* raw(short d) { return copies_[d].ptr; }
  void mark_written(short d)
  {
    for (auto& kv : copies_) kv.second.valid = (kv.first == d);
  }

 private:
  // Host copies are page-locked when a GPU is present (so copies on a stream run
  // asynchronously), plain heap memory otherwise; device copies come from the
  // engine's pool.  Fresh storage is zeroed.
  void allocate(short d, Copy& c)
  {
    const size_t bytes = n_ * sizeof(T);
    void* p = nullptr;
    if (IS_CPU_INDEX(d)) {
      int ndev = 0;
      if (nebo_device_count(&ndev) == NEBO_OK && ndev > 0 && nebo_host_alloc(bytes, &p) == NEBO_OK && p) {
        std::memset(p, 0, bytes);
        c.pinned = true;
        c.ptr = static_cast<T*>(p);
      } else {
        c.ptr = new T[n_]();
      }
    } else {
      nebo_check(nebo_alloc(d, bytes, &p));
      nebo_check(nebo_memset(p, 0, bytes, nullptr));
      c.ptr = static_cast<T*>(p);
    }
    c.owned = true;
  }
  void release(short d, const Copy& c)
  {
    if (!IS_CPU_INDEX(d)) nebo_free(d, c.ptr);
    else if (c.pinned) nebo_host_free(c.ptr);
    else delete[] c.ptr;
  }
  void copy(short dd, T* dst, short sd, const T* src, void* stream)
  {
    const size_t bytes = n_ * sizeof(T);
    if (IS_CPU_INDEX(dd) && IS_CPU_INDEX(sd)) std::memcpy(dst, src, bytes);
    else if (IS_CPU_INDEX(dd)) nebo_check(nebo_memcpy(dst, src, bytes, 1, stream));
    else if (IS_CPU_INDEX(sd)) nebo_check(nebo_memcpy(dst, src, bytes, 0, stream));
    else if (dd == sd) nebo_check(nebo_memcpy(dst, src, bytes, 2, stream));
    else nebo_check(nebo_memcpy_peer(dst, dd, src, sd, bytes, stream));
  }

  size_t n_;
  short active_;
  std::map<short, Copy> copies_;
  std::mutex mu_;
};

// ---------------------------------------------------------------------------
// SpatialField (SpatialField.h:128-603): a shallow handle; copies share storage.
// ---------------------------------------------------------------------------
template <typename Loc, typename T = double>
class SpatialField {
 public:
  typedef Loc Location;
  typedef T value_type;
  typedef SpatialField<Loc, T> field_type;

  SpatialField(const MemoryWindow& window, const BoundaryCellInfo& bc, const GhostData& ghosts,
               T* values = nullptr, StorageMode mode = InternalStorage, short devIdx = CPU_INDEX)
      : window_(window),
        bc_(bc.limit_by_extent(window.extent())),
        ghosts_(ghosts.limit_by_extent(window.extent())),
        valid_(ghosts_),
        info_(std::make_shared<FieldStorage<T>>(window.glob_npts(), values, mode, devIdx))
  {}

  // Copies share the storage (SpatialField.h:186-192); assignment copies the
  // values of the whole block from the other field's active device to this
  // field's active device (SpatialField.h:626-718).
  SpatialField(const SpatialField&) = default;
  SpatialField& operator=(const SpatialField& other)
  {
    if (window_ != other.window_)
      throw std::runtime_error("Error : Attempted assignment between fields of unequal size!");
    if (info_ == other.info_) return *this;
    const short dd = active_device_index(), sd = other.active_device_index();
    const size_t bytes = window_.glob_npts() * sizeof(T);
    T* dst = info_->raw(dd);
    const T* src = other.info_->const_values(sd);
    if (IS_CPU_INDEX(dd) && IS_CPU_INDEX(sd)) std::memcpy(dst, src, bytes);
    else if (IS_CPU_INDEX(dd)) nebo_check(nebo_memcpy(dst, src, bytes, 1, nullptr));
    else if (IS_CPU_INDEX(sd)) nebo_check(nebo_memcpy(dst, src, bytes, 0, nullptr));
    else if (dd == sd) nebo_check(nebo_memcpy(dst, src, bytes, 2, nullptr));
    else nebo_check(nebo_memcpy_peer(dst, dd, src, sd, bytes, nullptr));
    mark_written();
    return *this;
  }

  const MemoryWindow& window_with_ghost() const { return window_; }
  MemoryWindow window_without_ghost() const
  {
    return MemoryWindow(window_.glob_dim(), window_.offset() + ghosts_.get_minus(),
                        window_.extent() - ghosts_.get_minus() - ghosts_.get_plus() - bc_.has_extra());
  }
  const BoundaryCellInfo& boundary_info() const { return bc_; }
  const GhostData& get_ghost_data() const { return ghosts_; }
  const GhostData& get_valid_ghost_data() const { return valid_; }
  void reset_valid_ghosts(const GhostData& g) { valid_ = g; }
  size_t allocated_bytes() const { return window_.glob_npts() * sizeof(T); }

  short active_device_index() const { return info_->active(); }
  // the *_async forms queue the copy on the field's stream (set_stream) and
  // return; wait_for_synchronization() waits for it (SpatialField.h:282-446)
  void add_device(short d) { info_->add_device(d); }
  void add_device_async(short d) { info_->add_device(d, stream_); }
  void validate_device(short d) { info_->validate_device(d); }
  void validate_device_async(short d) { info_->validate_device(d, stream_); }
  void set_device_as_active(short d) { info_->set_device_as_active(d); }
  void set_device_as_active_async(short d) { info_->set_device_as_active(d, stream_); }
  bool is_valid(short d) const { return info_->is_valid(d); }
  bool is_available(short d) const { return info_->is_available(d); }
  T* field_values(short d = CPU_INDEX) { return info_->values(d); }
  const T* field_values(short d = CPU_INDEX) const { return info_->const_values(d); }

  void set_stream(void* s) { stream_ = s; }
  void* get_stream() const { return stream_; }
  void wait_for_synchronization() { nebo_check(nebo_stream_sync(stream_)); }

  // host element access (the active copy must be the CPU one)
  T& operator()(int i, int j, int k) { return field_values(CPU_INDEX)[window_.flat_index(IntVec(i, j, k))]; }
  const T& operator()(int i, int j, int k) const
  {
    return field_values(CPU_INDEX)[window_.flat_index(IntVec(i, j, k))];
  }
  T& operator[](size_t i) { return field_values(CPU_INDEX)[i]; }

  // host iterators over the window with ghosts and over the interior
  // (SpatialField.h:448-499); the CPU copy must be the active one
  typedef WindowIterator<T> iterator;
  typedef WindowIterator<const T> const_iterator;
  iterator begin() { return iterator(field_values(CPU_INDEX), window_); }
  iterator end() { return begin() + window_.local_npts(); }
  const_iterator begin() const { return const_iterator(field_values(CPU_INDEX), window_); }
  const_iterator end() const { return begin() + window_.local_npts(); }
  iterator interior_begin() { return iterator(field_values(CPU_INDEX), window_without_ghost()); }
  iterator interior_end() { return interior_begin() + window_without_ghost().local_npts(); }

  // the engine-side description of this field on its active device
  nebo_field_t descriptor() const
  {
    nebo_field_t f;
    std::memset(&f, 0, sizeof f);
    const short d = info_->active();
    f.values = info_->raw(d);
    for (int a = 0; a < 3; ++a) {
      f.glob[a] = window_.glob_dim()[a];
      f.offset[a] = window_.offset()[a];
      f.extent[a] = window_.extent()[a];
      f.ghost_minus[a] = ghosts_.get_minus(a);
      f.ghost_plus[a] = ghosts_.get_plus(a);
      f.valid_minus[a] = valid_.get_minus(a);
      f.valid_plus[a] = valid_.get_plus(a);
      f.has_bc[a] = bc_.has_bc(a);
      f.bc_extra[a] = bc_.has_extra(a);
    }
    f.device = d;
    return f;
  }
  const void* storage_id() const { return info_.get(); }
  void mark_written() { info_->mark_written(info_->active()); }

 private:
  MemoryWindow window_;
  BoundaryCellInfo bc_;
  GhostData ghosts_;
  GhostData valid_;
  std::shared_ptr<FieldStorage<T>> info_;
  void* stream_ = nullptr;
};

// A window / boundary info for a field of type FieldT next to a prototype of
// another type: the prototype's window grown by the difference of the two
// locations' BC extra cells on the axes that carry a boundary
// (create_new_memory_window / create_new_boundary_cell_info, SpatialField.h:720-755).
template <typename FieldT, typename ProtoT>
BoundaryCellInfo create_new_boundary_cell_info(const ProtoT& prototype)
{
  return BoundaryCellInfo::build<FieldT>(prototype.boundary_info().has_bc());
}

template <typename FieldT, typename ProtoT>
MemoryWindow create_new_memory_window(const ProtoT& prototype)
{
  const BoundaryCellInfo bc = create_new_boundary_cell_info<FieldT, ProtoT>(prototype);
  const MemoryWindow& w = prototype.window_with_ghost();
  const IntVec inc = bc.has_bc() * (FieldT::Location::bc_extra() - ProtoT::Location::bc_extra());
  return MemoryWindow(w.glob_dim() + inc, w.offset(), w.extent() + inc);
}

// field typedefs (FVStaggeredFieldTypes.h)
typedef SpatialField<SVol> SVolField;
typedef SpatialField<SSurfX> SSurfXField;
typedef SpatialField<SSurfY> SSurfYField;
typedef SpatialField<SSurfZ> SSurfZField;
typedef SpatialField<XVol> XVolField;
typedef SpatialField<XSurfX> XSurfXField;
typedef SpatialField<XSurfY> XSurfYField;
typedef SpatialField<XSurfZ> XSurfZField;
typedef SpatialField<YVol> YVolField;
typedef SpatialField<YSurfX> YSurfXField;
typedef SpatialField<YSurfY> YSurfYField;
typedef SpatialField<YSurfZ> YSurfZField;
typedef SpatialField<ZVol> ZVolField;
typedef SpatialField<ZSurfX> ZSurfXField;
typedef SpatialField<ZSurfY> ZSurfYField;
typedef SpatialField<ZSurfZ> ZSurfZField;
typedef SpatialField<SingleValue> SingleValueField;

template <typename VolT> struct FaceTypes;
template <> struct FaceTypes<SVolField> { typedef SSurfXField XFace; typedef SSurfYField YFace; typedef SSurfZField ZFace; };
template <> struct FaceTypes<XVolField> { typedef XSurfXField XFace; typedef XSurfYField YFace; typedef XSurfZField ZFace; };
template <> struct FaceTypes<YVolField> { typedef YSurfXField XFace; typedef YSurfYField YFace; typedef YSurfZField ZFace; };
template <> struct FaceTypes<ZVolField> { typedef ZSurfXField XFace; typedef ZSurfYField YFace; typedef ZSurfZField ZFace; };

}  // namespace SpatialOps
