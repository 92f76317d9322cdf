// simple_pcd_loader.h -- PCD load / save (src/io/simple_pcd_loader.h:143-328) over the library's reader
// (ascii, binary and LZF binary_compressed; x/y/z of any record layout) and writer.
#pragma once

#include <string>
#include <vector>

#include "core/point_types.h"

namespace rvpoint {

// Returns the number of points read, or -1 on error (the reference's convention).
int loadPCD(const std::string &file_path, std::vector<PointXYZ> &points);

// DATA ascii (default stream formatting) or binary, FIELDS x y z.
void savePCD(const std::string &filename, const std::vector<PointXYZ> &points, bool binary = false);

}  // namespace rvpoint
