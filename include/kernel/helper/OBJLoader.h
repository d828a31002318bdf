// Wavefront OBJ mesh reader of the LSSolver drop-in (FunDEM/kernel/helper/OBJLoader.h:133-282),
// written anew: `v x y z` records give vertices, `f` records faces (vertex index of each token,
// before any '/'; negative indices count back from the last vertex; polygons are split into a fan
// of triangles around their first vertex).  Other records are skipped.  A relative file name is
// resolved against the directory of the executable (argv[0]), as the reference apps expect.
#pragma once

#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../lsdem_types.h"

namespace OBJLoader {

namespace detail {
inline std::string resolve(const std::string& file, const char* argv0) {
  const std::filesystem::path p(file);
  if (p.is_absolute()) return p.string();
  std::filesystem::path dir = std::filesystem::current_path();
  if (argv0 != nullptr) dir = std::filesystem::weakly_canonical(std::filesystem::absolute(argv0)).parent_path();
  return (dir / p).string();
}

inline std::string strip(const std::string& s) {
  const auto a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const auto b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// Zero-based vertex index of a face token ("i", "i/t", "i//n", "i/t/n").
inline int faceIndex(const std::string& token, int nVertices) {
  const std::string head = token.substr(0, token.find('/'));
  if (head.empty()) throw std::runtime_error("Invalid OBJ face token: " + token);
  const int i = std::stoi(head);
  if (i == 0) throw std::runtime_error("OBJ index cannot be zero.");
  const int z = i > 0 ? i - 1 : nVertices + i;
  if (z < 0 || z >= nVertices) throw std::runtime_error("Face index out of range");
  return z;
}
}  // namespace detail

// Reads vertices and triangles; returns false (lists cleared, message on stderr) when the file
// cannot be opened or a record is malformed.
inline bool loadOBJMesh(const std::string& filename, std::vector<double3>& vertices, std::vector<int3>& faces,
                        const int argc, char** argv, const bool verbose = true) {
  vertices.clear();
  faces.clear();
  const std::string path = detail::resolve(filename, (argc > 0 && argv != nullptr) ? argv[0] : nullptr);
  std::ifstream in(path);
  if (!in.is_open()) {
    std::cerr << "[OBJLoader] Failed to open file: " << path << std::endl;
    return false;
  }
  std::string line;
  size_t lineNo = 0;
  try {
    while (std::getline(in, line)) {
      ++lineNo;
      line = detail::strip(line);
      if (line.empty() || line[0] == '#') continue;
      std::istringstream rec(line);
      std::string tag;
      rec >> tag;
      if (tag == "v") {
        double x, y, z;
        rec >> x >> y >> z;
        if (rec.fail()) throw std::runtime_error("Failed to parse vertex");
        vertices.push_back(make_double3(x, y, z));
      } else if (tag == "f") {
        std::vector<int> poly;
        std::string tok;
        while (rec >> tok) poly.push_back(detail::faceIndex(tok, int(vertices.size())));
        if (poly.size() < 3) throw std::runtime_error("Face has fewer than 3 vertices");
        for (size_t k = 1; k + 1 < poly.size(); ++k) faces.push_back(make_int3(poly[0], poly[k], poly[k + 1]));
      }
    }
  } catch (const std::exception& e) {
    std::cerr << "[OBJLoader] Error while reading " << path << " at line " << lineNo << ": " << e.what()
              << std::endl;
    vertices.clear();
    faces.clear();
    return false;
  }
  if (verbose)
    std::cout << "[OBJLoader] Loaded " << path << ": " << vertices.size() << " vertices, " << faces.size()
              << " triangles" << std::endl;
  return true;
}

}  // namespace OBJLoader
