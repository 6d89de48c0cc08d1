// SpatialMask.h -- SpatialMask<FieldT> of the C++ front end (defined with the
// expression language, which it feeds as a boolean leaf).
#pragma once
#include <spatialops/Nebo.h>
