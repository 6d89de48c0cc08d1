// Temporary fields and field residency of the C++ front end:
//  * test_store of spatialops/structured/test/testField.cpp: temporaries from
//    SpatialFieldStore::get / get_from_window have the windows the reference
//    derives for every location pair, with and without boundaries;
//  * field assignment copies values (the "spatial field pointer from store"
//    check of testField.cpp), between host and GPU copies too;
//  * GPU temporaries come from the engine's pool: releasing the last SpatFldPtr
//    returns the block, and the next temporary of that size reuses it;
//  * validity tracking: a GPU write makes the host copy stale until the field
//    is moved back;
//  * asynchronous residency: copies queued on the field's stream with the
//    *_async calls, ordered with the assignments on that stream.
#include <spatialops/structured/FVStaggered.h>

#include <hip/hip_runtime_api.h>

#include <sstream>

#include "nebo_test_helper.h"

using namespace SpatialOps;

template <typename FT1, typename FT2>
bool test_store(const IntVec& dim, const IntVec& bc, short dev)
{
  TestHelper status;
  const GhostData ghost1(1), ghost2(1);
  const BoundaryCellInfo bc1 = BoundaryCellInfo::build<FT1>(bc[0], bc[1], bc[2]);
  const BoundaryCellInfo bc2 = BoundaryCellInfo::build<FT2>(bc[0], bc[1], bc[2]);
  const MemoryWindow w1 = get_window_with_ghost(dim, ghost1, bc1);
  const MemoryWindow w2 = get_window_with_ghost(dim, ghost2, bc2);
  SpatFldPtr<FT1> f1 = SpatialFieldStore::get_from_window<FT1>(w1, bc1, ghost1, dev);
  SpatFldPtr<FT2> f2 = SpatialFieldStore::get_from_window<FT2>(w2, bc2, ghost2, dev);
  SpatFldPtr<FT1> f1a = SpatialFieldStore::get<FT1>(*f1);
  SpatFldPtr<FT2> f2a = SpatialFieldStore::get<FT2>(*f2);
  SpatFldPtr<FT2> f2b = SpatialFieldStore::get<FT2>(*f1);
  status(f1->window_with_ghost() == f1a->window_with_ghost(), "f1==f1a");
  status(f2->window_with_ghost() == f2a->window_with_ghost(), "f2==f2a");
  status(f2->window_with_ghost() == f2b->window_with_ghost(), "f2==f2b");
  status(f2b->active_device_index() == dev && f2b.built_from_store(), "on the prototype's device");
  return status.ok;
}

int main()
{
  TestHelper status;
  for (short dev : {(short)CPU_INDEX, (short)GPU_INDEX}) {
    const std::string where = IS_CPU_INDEX(dev) ? " (CPU)" : " (GPU)";
    const IntVec d(30, 40, 50), none(0, 0, 0), all(1, 1, 1);
    status(test_store<SVolField, SVolField>(d, none, dev), "SVol,SVol store" + where);
    status(test_store<SVolField, SVolField>(d, all, dev), "SVol,SVol(bc) store" + where);
    status(test_store<SVolField, SSurfXField>(d, none, dev), "SVol,SSX store" + where);
    status(test_store<SSurfXField, SVolField>(d, none, dev), "SSX,SVol store" + where);
    status(test_store<SVolField, SSurfYField>(d, all, dev), "SVol,SSY(bc) store" + where);
    status(test_store<SVolField, SSurfYField>(d, none, dev), "SVol,SSY store" + where);
    status(test_store<SVolField, SSurfXField>(d, all, dev), "SVol,SSX(bc) store" + where);
    status(test_store<SSurfXField, SVolField>(d, all, dev), "SSX,SVol(bc) store" + where);
    status(test_store<SVolField, SSurfZField>(d, all, dev), "SVol,SSZ(bc) store" + where);
    status(test_store<SVolField, SSurfZField>(d, none, dev), "SVol,SSZ store" + where);
    status(test_store<XSurfXField, SVolField>(d, none, dev), "XSX,SVol store" + where);
    status(test_store<XSurfXField, ZVolField>(d, none, dev), "XSX,ZVol store" + where);
    status(test_store<YSurfZField, ZSurfXField>(d, none, dev), "YSZ,ZSX store" + where);
  }

  // field operations and assignment from a store temporary (testField.cpp:500-535)
  {
    const IntVec n(5, 6, 7);
    const GhostData g(1);
    const BoundaryCellInfo bc = BoundaryCellInfo::build<SVolField>(false, false, false);
    const MemoryWindow w = get_window_with_ghost(n, g, bc);
    SVolField svol1(w, bc, g, nullptr, InternalStorage, CPU_INDEX);
    for (int k = 0; k < (int)w.extent(2); ++k)
      for (int j = 0; j < (int)w.extent(1); ++j)
        for (int i = 0; i < (int)w.extent(0); ++i) svol1(i, j, k) = i + j + k;
    svol1.set_device_as_active(GPU_INDEX);
    SpatFldPtr<SVolField> svol2 = SpatialFieldStore::get<SVolField>(svol1);
    *svol2 <<= 2.0;
    svol1 <<= (svol1 + *svol2) * *svol2 / *svol2 - *svol2;
    {
      SpatFldPtr<SVolField> sv3 = SpatialFieldStore::get<SVolField>(svol1, CPU_INDEX);
      *sv3 = svol1;   // GPU -> CPU values
      bool same = true;
      for (int k = 0; k < (int)w.extent(2); ++k)
        for (int j = 0; j < (int)w.extent(1); ++j)
          for (int i = 0; i < (int)w.extent(0); ++i) same = same && (*sv3)(i, j, k) == i + j + k;
      status(same, "spatial field pointer from store (GPU field assigned to a CPU temporary)");
      status(sv3.count() == 1 && svol2.count() == 1, "SpatFldPtr counts");
      SpatFldPtr<SVolField> alias = sv3;
      status(sv3.count() == 2, "SpatFldPtr copies share the field");
    }
    // validity: the GPU write left the host copy stale
    status(!svol1.is_valid(CPU_INDEX) && svol1.is_valid(GPU_INDEX), "host copy stale after a GPU write");
    svol1.set_device_as_active(CPU_INDEX);
    status(svol1.is_valid(CPU_INDEX) && svol1(2, 3, 4) == 9.0, "moved back to the host");
  }

  // asynchronous moves on the field's stream, ordered with its assignments
  {
    const IntVec n(40, 30, 20);
    const GhostData g(1);
    const BoundaryCellInfo bc = BoundaryCellInfo::build<SVolField>(false, false, false);
    const MemoryWindow w = get_window_with_ghost(n, g, bc);
    hipStream_t s = nullptr;
    status(hipStreamCreate(&s) == hipSuccess, "stream created");
    SVolField f(w, bc, g, nullptr, InternalStorage, CPU_INDEX);
    f.set_stream(s);
    for (size_t i = 0; i < w.glob_npts(); ++i) f[i] = double(i % 1000);
    for (int rep = 0; rep < 3; ++rep) {
      f.set_device_as_active_async(GPU_INDEX);   // queued host -> device copy
      f <<= f * 2.0 + 1.0;                        // same stream: runs after the copy
      f.set_device_as_active_async(CPU_INDEX);   // queued device -> host copy
      f.wait_for_synchronization();
    }
    bool same = true;
    for (size_t i = 0; i < w.glob_npts(); ++i) {
      double v = double(i % 1000);
      for (int rep = 0; rep < 3; ++rep) v = v * 2.0 + 1.0;
      same = same && f[i] == v;
    }
    status(same, "async moves and assignments on one stream");
    status(hipStreamDestroy(s) == hipSuccess, "stream destroyed");
  }

  // GPU temporaries reuse pooled blocks
  {
    const IntVec n(64, 64, 64);
    const GhostData g(1);
    const BoundaryCellInfo bc = BoundaryCellInfo::build<SVolField>(false, false, false);
    const MemoryWindow w = get_window_with_ghost(n, g, bc);
    size_t used0 = 0, high0 = 0, used1 = 0, high1 = 0, used2 = 0, high2 = 0;
    nebo_check(nebo_pool_stats(GPU_INDEX, &used0, &high0));
    {
      SpatFldPtr<SVolField> a = SpatialFieldStore::get_from_window<SVolField>(w, bc, g, GPU_INDEX);
      *a <<= 1.0;
      nebo_check(nebo_pool_stats(GPU_INDEX, &used1, &high1));
    }
    for (int rep = 0; rep < 5; ++rep) {
      SpatFldPtr<SVolField> b = SpatialFieldStore::get_from_window<SVolField>(w, bc, g, GPU_INDEX);
      *b <<= 2.0;
      status(nebo_sum(*b) == 2.0 * w.glob_npts(), "temporary evaluates");
    }
    nebo_check(nebo_pool_stats(GPU_INDEX, &used2, &high2));
    std::ostringstream m;
    m << "pool reuse (in use " << used0 << " -> " << used1 << " -> " << used2 << ", high water " << high1
      << " -> " << high2 << ")";
    status(used1 >= used0 + w.glob_npts() * sizeof(double) && used2 == used0 && high2 == high1, m.str());
  }
  return status.report();
}
