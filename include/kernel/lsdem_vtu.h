// VTK XML unstructured-grid writer used by the LSSolver drop-in for its frame output.
//
// Files follow the layout the reference's writers produce (LSParticle.h:311-654,
// SolidInteraction.h:154-317, VBondedInteraction.h:262-448): inline binary arrays, each
// base64( [UInt64 byte count][little-endian data] ), header_type UInt64, FieldData TIME
// (Float32) and STEP (Int32), Float32 points.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace lsdem_vtu {

struct Array {
  std::string name;     // empty: no Name attribute (Points)
  const char* type;     // "Float32", "Int32", "UInt8"
  int components = 0;   // 0: no NumberOfComponents attribute
  std::vector<uint8_t> bytes;
};

template <class T>
inline Array array(const std::string& name, const char* type, int components, const std::vector<T>& v) {
  Array a;
  a.name = name;
  a.type = type;
  a.components = components;
  a.bytes.resize(v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(a.bytes.data(), v.data(), a.bytes.size());
  return a;
}

inline void putBase64(std::ostream& out, const std::vector<uint8_t>& payload) {
  static const char table[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::vector<uint8_t> buf(8 + payload.size());
  const uint64_t n = payload.size();
  std::memcpy(buf.data(), &n, 8);
  if (!payload.empty()) std::memcpy(buf.data() + 8, payload.data(), payload.size());
  std::string s;
  s.reserve((buf.size() + 2) / 3 * 4);
  for (size_t i = 0; i < buf.size(); i += 3) {
    const size_t left = buf.size() - i;
    const uint32_t w = (uint32_t(buf[i]) << 16) | (left > 1 ? uint32_t(buf[i + 1]) << 8 : 0u) |
                       (left > 2 ? uint32_t(buf[i + 2]) : 0u);
    s.push_back(table[(w >> 18) & 63]);
    s.push_back(table[(w >> 12) & 63]);
    s.push_back(left > 1 ? table[(w >> 6) & 63] : '=');
    s.push_back(left > 2 ? table[w & 63] : '=');
  }
  out << s;
}

inline void putArray(std::ostream& out, const Array& a) {
  out << "<DataArray type=\"" << a.type << "\"";
  if (!a.name.empty()) out << " Name=\"" << a.name << "\"";
  if (a.components > 0) out << " NumberOfComponents=\"" << a.components << "\"";
  out << " format=\"binary\">\n";
  putBase64(out, a.bytes);
  out << "\n</DataArray>\n";
}

// One piece: points (Float32 x 3), optional cells (connectivity/offsets/types), point data.
// `vectorsAttr` / `scalarsAttr` name the active point arrays (may be empty).
inline void write(const std::string& file, double time, size_t step, const std::vector<float>& points,
                  const std::vector<int32_t>* connectivity, const std::vector<int32_t>* offsets,
                  const std::vector<uint8_t>* types, const std::vector<Array>& pointData,
                  const std::string& vectorsAttr = "", const std::string& scalarsAttr = "") {
  std::ofstream out(file);
  if (!out) throw std::runtime_error("Cannot open for writing: " + file);
  const size_t nPoints = points.size() / 3;
  const size_t nCells = types ? types->size() : 0;
  out << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
      << "<UnstructuredGrid>\n"
      << "<FieldData>\n"
      << "<DataArray type=\"Float32\" Name=\"TIME\" NumberOfTuples=\"1\" format=\"ascii\">" << float(time)
      << "</DataArray>\n"
      << "<DataArray type=\"Int32\" Name=\"STEP\" NumberOfTuples=\"1\" format=\"ascii\">" << int32_t(step)
      << "</DataArray>\n"
      << "</FieldData>\n"
      << "<Piece NumberOfPoints=\"" << nPoints << "\" NumberOfCells=\"" << nCells << "\">\n"
      << "<Points>\n";
  putArray(out, array("", "Float32", 3, points));
  out << "</Points>\n";
  if (nCells) {
    out << "<Cells>\n";
    putArray(out, array("connectivity", "Int32", 0, *connectivity));
    putArray(out, array("offsets", "Int32", 0, *offsets));
    putArray(out, array("types", "UInt8", 0, *types));
    out << "</Cells>\n";
  } else {
    out << "<Cells/>\n";
  }
  out << "<PointData";
  if (!vectorsAttr.empty()) out << " Vectors=\"" << vectorsAttr << "\"";
  if (!scalarsAttr.empty()) out << " Scalars=\"" << scalarsAttr << "\"";
  out << ">\n";
  for (const Array& a : pointData) putArray(out, a);
  out << "</PointData>\n"
      << "</Piece>\n"
      << "</UnstructuredGrid>\n"
      << "</VTKFile>\n";
  if (!out) throw std::runtime_error("Write error on: " + file);
}

// Vertex cells (VTK_VERTEX = 1), one per point.
inline void vertexCells(size_t n, std::vector<int32_t>& conn, std::vector<int32_t>& offs, std::vector<uint8_t>& types) {
  conn.resize(n);
  offs.resize(n);
  types.assign(n, uint8_t(1));
  for (size_t i = 0; i < n; ++i) {
    conn[i] = int32_t(i);
    offs[i] = int32_t(i + 1);
  }
}

// ParaView collection of the frames of one series.
inline void writePvd(const std::string& file, const std::vector<std::pair<double, std::string>>& frames) {
  std::ofstream out(file);
  if (!out) return;
  out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << "<Collection>\n";
  for (const auto& f : frames)
    out << "<DataSet timestep=\"" << f.first << "\" group=\"\" part=\"0\" file=\"" << f.second << "\"/>\n";
  out << "</Collection>\n</VTKFile>\n";
}

}  // namespace lsdem_vtu
