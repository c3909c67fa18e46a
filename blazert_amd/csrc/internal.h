// Library-internal hooks shared by scene.cpp and autotune.cpp (not part of the C ABI).
#pragma once
#include <string>

#include "../../include/blazert_amd.h"

namespace brt {

// Records `msg` as the calling thread's last error (brt_last_error) and returns `code`.
int fail(int code, const std::string &msg);
// Builds the scene again with new build options (the scene's copied geometry is kept) and drops its device copy;
// the scene is committed afterwards.
int scene_rebuild(brt_scene *scene, const brt_build_options &opts);
brt_build_options scene_options(const brt_scene *scene);
// Launch configuration of a batch, and its replacement (refused once the batch has been captured in a graph).
brt_launch batch_launch_config(const brt_batch *batch);
int batch_set_launch_config(brt_batch *batch, const brt_launch &launch);
// Validates the tunable fields of a launch (precision, block size, refill, stack ring, sort, schedule).
int check_launch(const brt_launch &launch);

}  // namespace brt
