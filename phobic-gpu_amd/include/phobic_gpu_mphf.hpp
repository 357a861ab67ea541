// phobic_gpu_mphf.hpp -- umbrella header of the MI355X PHOBIC build path,
// the same entry point as the reference's include/phobic_gpu_mphf.hpp:1-22.
#pragma once

#include "encoders/bit_vector.hpp"
#include "encoders/compact_vector.hpp"
#include "encoders/rice_sequence.hpp"
#include "encoders/diff_partition_offset_encoder.hpp"
#include "encoders/direct_partition_offset_encoder.hpp"
#include "encoders/encoders.hpp"
#include "encoders/interleaved_encoder.hpp"
#include "encoders/interleaved_encoder_dual.hpp"
#include "encoders/mono_encoder.hpp"
#include "phobicGpu/bucketers.h"
#include "phobicGpu/hasher.hpp"
#include "phobicGpu/host_timer.h"
#include "phobicGpu/mphf.hpp"
#include "phobicGpu/mphf_builder.h"
#include "phobicGpu/mphf_config.h"

namespace phobicgpu {

typedef MPHF<interleaved_encoder<compact>, diff_partition_encoder<compact>, xxhash> FastQueryMphf;
typedef MPHF<interleaved_encoder<rice>, diff_partition_encoder<compact>, xxhash> SmallSpaceMphf;

}  // namespace phobicgpu
