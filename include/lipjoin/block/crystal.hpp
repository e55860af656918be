// crystal.hpp -- umbrella header of the wave64 block operator API; includes
// the same groups of functions as crystal/crystal.cuh:25-32.
#pragma once

#include "load.hpp"
#include "pred.hpp"
#include "store.hpp"
#include "reduce.hpp"
#include "join.hpp"
#include "term.hpp"
#include "bloom.hpp"

// Rows per block-wide tile, as gpu_utils.h:47 defines it for the drivers.
#ifndef TILE_SIZE
#define TILE_SIZE (BLOCK_THREADS * ITEMS_PER_THREAD)
#endif
