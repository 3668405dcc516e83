// Host-side model: the JSON scalars and the twelve float32 input arrays.
#pragma once

#include <string>
#include <vector>

#include "msfdm.h"

struct msfdm_model {
    msfdm_desc d{};
    std::vector<int> src_x, src_z;
    std::vector<float> a[MSFDM_NARRAY];
};

namespace msfdm {

// Thread-local error reporting shared by every C entry point.
int set_error(int code, const std::string &msg);
void clear_error();

// Rows of model array `which` (fluid, solid or global grid rows).
int array_rows(const msfdm_desc &d, int which);

// Rows a source row index src_z refers to: the fluid array of a fluid-solid
// grid, the solid array of a solid-only grid.
int source_rows(const msfdm_desc &d);

// Validate the scalar description (grid sizes, PML fits, supported modes).
// Returns MSFDM_OK or sets the error.
int validate_desc(const msfdm_desc &d);

// Record interval in steps: (int)(1e-3 / dt + 0.5), src/main.cu:92.
int record_gap(float dt);

}  // namespace msfdm
