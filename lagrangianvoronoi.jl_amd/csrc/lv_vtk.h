// lv_vtk.h -- VTK XML PolyData (.vtp) writer for export_grid / export_points (src/IO.jl:9-71).
#pragma once

#include <stdint.h>

// one cell-data array: n values of ncomp components (1 or 3), interleaved
struct lv_vtk_array {
    const char *name;
    int ncomp;
    const double *data;
};

enum { LV_VTK_POLYS = 0, LV_VTK_VERTS = 1 };

// Writes path (PolyData, one piece): npoints points (x, y, z interleaved), ncells cells of the given
// kind with Int64 connectivity (0-based) and offsets (end of each cell), and the cell-data arrays.
// Data go in the appended section in raw binary, each array zlib-compressed as one block when
// compress is set.  Returns 0, or -1 when the file cannot be written.
int lv_write_vtp(const char *path, int64_t npoints, const double *points3, int kind, int64_t ncells,
                 const int64_t *connectivity, const int64_t *offsets, const lv_vtk_array *arrays, int narrays,
                 int compress);
