// lv_vtk.hip -- host code: the VTK XML PolyData files that export_grid / export_points
// (src/IO.jl:9-71) write through WriteVTK.jl 1.18.3 (Manifest.toml:1183): one piece, the geometry
// and cell data in the appended section in raw binary with UInt64 headers, each array compressed
// with zlib as a single block (WriteVTK's default) or stored as is.
#include "lv_vtk.h"

#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <string>
#include <vector>

namespace {

struct blob {
    std::vector<unsigned char> bytes;  // header + payload, as it goes into the appended section
};

// the appended representation of nbytes of data: [nbytes] raw, or [1, nbytes, nbytes, csize] and
// the zlib stream
bool make_blob(const void *data, uint64_t nbytes, int compress, blob &b)
{
    b.bytes.clear();
    auto put64 = [&](uint64_t v) {
        const unsigned char *p = (const unsigned char *)&v;
        b.bytes.insert(b.bytes.end(), p, p + 8);
    };
    if (!compress) {
        put64(nbytes);
        const unsigned char *p = (const unsigned char *)data;
        b.bytes.insert(b.bytes.end(), p, p + nbytes);
        return true;
    }
    uLongf clen = compressBound((uLong)nbytes);
    std::vector<unsigned char> out(clen);
    if (compress2(out.data(), &clen, (const Bytef *)data, (uLong)nbytes, Z_DEFAULT_COMPRESSION) != Z_OK) return false;
    put64(1);
    put64(nbytes);
    put64(nbytes);
    put64((uint64_t)clen);
    b.bytes.insert(b.bytes.end(), out.begin(), out.begin() + clen);
    return true;
}

std::string xml_escape(const char *s)
{
    std::string o;
    for (; *s; s++) {
        switch (*s) {
        case '&': o += "&amp;"; break;
        case '<': o += "&lt;"; break;
        case '>': o += "&gt;"; break;
        case '"': o += "&quot;"; break;
        default: o += *s;
        }
    }
    return o;
}

}  // namespace

int lv_write_vtp(const char *path, int64_t npoints, const double *points3, int kind, int64_t ncells,
                 const int64_t *connectivity, const int64_t *offsets, const lv_vtk_array *arrays, int narrays,
                 int compress)
{
    const int64_t nconn = ncells > 0 ? offsets[ncells - 1] : 0;
    std::vector<blob> blobs(3 + narrays);
    if (!make_blob(points3, (uint64_t)npoints * 3 * sizeof(double), compress, blobs[0]) ||
        !make_blob(connectivity, (uint64_t)nconn * sizeof(int64_t), compress, blobs[1]) ||
        !make_blob(offsets, (uint64_t)ncells * sizeof(int64_t), compress, blobs[2]))
        return -1;
    for (int a = 0; a < narrays; a++)
        if (!make_blob(arrays[a].data, (uint64_t)ncells * arrays[a].ncomp * sizeof(double), compress, blobs[3 + a]))
            return -1;
    std::vector<uint64_t> off(blobs.size());
    uint64_t o = 0;
    for (size_t k = 0; k < blobs.size(); k++) {
        off[k] = o;
        o += blobs[k].bytes.size();
    }
    const char *cells = (kind == LV_VTK_VERTS) ? "Verts" : "Polys";
    std::string x;
    x += "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n";
    x += "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\"";
    if (compress) x += " compressor=\"vtkZLibDataCompressor\"";
    x += ">\n  <PolyData>\n";
    x += "    <Piece NumberOfPoints=\"" + std::to_string(npoints) + "\" NumberOfVerts=\"" +
         std::to_string(kind == LV_VTK_VERTS ? ncells : 0) + "\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"" +
         std::to_string(kind == LV_VTK_POLYS ? ncells : 0) + "\">\n";
    x += "      <Points>\n        <DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\" format=\"appended\" "
         "offset=\"" + std::to_string(off[0]) + "\"/>\n      </Points>\n";
    x += std::string("      <") + cells + ">\n";
    x += "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"appended\" offset=\"" + std::to_string(off[1]) +
         "\"/>\n";
    x += "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" offset=\"" + std::to_string(off[2]) +
         "\"/>\n";
    x += std::string("      </") + cells + ">\n";
    if (narrays > 0) {
        x += "      <CellData>\n";
        for (int a = 0; a < narrays; a++) {
            x += "        <DataArray type=\"Float64\" Name=\"" + xml_escape(arrays[a].name) + "\"";
            if (arrays[a].ncomp != 1) x += " NumberOfComponents=\"" + std::to_string(arrays[a].ncomp) + "\"";
            x += " format=\"appended\" offset=\"" + std::to_string(off[3 + a]) + "\"/>\n";
        }
        x += "      </CellData>\n";
    }
    x += "    </Piece>\n  </PolyData>\n  <AppendedData encoding=\"raw\">\n_";
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    bool ok = fwrite(x.data(), 1, x.size(), f) == x.size();
    for (const blob &b : blobs) ok = ok && fwrite(b.bytes.data(), 1, b.bytes.size(), f) == b.bytes.size();
    const char *tail = "\n  </AppendedData>\n</VTKFile>\n";
    ok = ok && fwrite(tail, 1, strlen(tail), f) == strlen(tail);
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}
