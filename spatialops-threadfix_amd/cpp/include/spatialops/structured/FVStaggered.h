// FVStaggered.h -- field types, operator types and the expression language.
#pragma once
#include <spatialops/structured/SpatialField.h>
#include <spatialops/structured/SpatialFieldStore.h>
#include <spatialops/Nebo.h>
#include <spatialops/NeboStencilBuilder.h>
