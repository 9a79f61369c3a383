// Output of the `mif` driver (include/VTKDatExport.h, src/VTKDatExport.cpp
// of the reference), for one rank holding the whole domain.  The values are
// fetched from the device point by point (mif_field_gather), so only the
// printed points cross PCIe.
#ifndef MIF_SHIM_VTK_DAT_EXPORT_H
#define MIF_SHIM_VTK_DAT_EXPORT_H

#include <string>

#include "VelocityTensor.h"

namespace mif {

// The planes x = 0, z = 0 and y = 0 (in that order in the file: z first) as
// a legacy binary VTK unstructured grid of points: coordinates, then u, v,
// w interpolated to the pressure points and p, big-endian (src/VTKDatExport.
// cpp:115-313).
void writeVTK(const std::string &filename, const VelocityTensor &velocity, const StaggeredTensor &pressure);

// A line parallel to axis `direction` (0 = x, 1 = y, 2 = z) through the
// point (x, y, z), which must lie on a pressure plane or halfway between two
// on each other axis: "x y z u v w p" per point, sorted by the running
// coordinate (src/VTKDatExport.cpp:342-584).
void writeDat(const std::string &filename, const VelocityTensor &velocity, const StaggeredTensor &pressure,
              const int direction, const Real x, const Real y, const Real z);

// The whole mesh as an ASCII VTK structured-points file with u, v, w, |u|
// and p at the pressure points (src/VTKDatExport.cpp:587-669).
void writeVTKFullMesh(const std::string &filename, const VelocityTensor &velocity, const StaggeredTensor &pressure);

}  // namespace mif

#endif  // MIF_SHIM_VTK_DAT_EXPORT_H
