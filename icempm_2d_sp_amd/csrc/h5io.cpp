// h5io.cpp — the reference's HDF5 formats around the substep, host side only
// (HDF5 1.10 C++ API): snapshots (SnapshotManager::SaveSnapshot /
// ReadPointsFromSnapshot, simulation/data_manager/snapshotmanager.cpp:308-396)
// and the static current field (WindAndCurrentInterpolator::ReadStaticFlowData,
// simulation/data_manager/windandcurrentinterpolator.cpp:18-63).
#include "icempm_io.h"

#include <H5Cpp.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kRefLines = 19;   // SimParams::nPtsArrays
constexpr int kIntLines = 3;    // utility, cell, point index: bit patterns
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

uint32_t float_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

float bits_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// An integer in a double slot the reference's way: the low 32 bits of the
// double's bit pattern (a reinterpret_cast to uint32_t* on a little-endian
// host), the high bits zero.
double int_in_double(uint32_t u) {
    const uint64_t b = u;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

uint32_t int_from_double(double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return (uint32_t)(b & 0xffffffffu);
}

template <class T>
void write_scalar_attr(H5::DataSet &ds, const char *name, const H5::PredType &t, const T &v) {
    H5::DataSpace scalar(H5S_SCALAR);
    ds.createAttribute(name, t, scalar).write(t, &v);
}

H5::DSetCreatPropList chunked(int rank, const hsize_t *chunk) {
    H5::DSetCreatPropList pl;
    pl.setChunk(rank, chunk);
    pl.setDeflate(5);
    return pl;
}

struct Snap {
    uint32_t n = 0;
    int step = 0, lines = 0;
    double time = 0.0;
    size_t elem = 0;
    bool cohesion = false;
};

Snap read_header(H5::H5File &f, H5::DataSet &ds) {
    Snap s;
    ds.openAttribute("SimulationStep").read(H5::PredType::NATIVE_INT, &s.step);
    ds.openAttribute("SimulationTime").read(H5::PredType::NATIVE_DOUBLE, &s.time);
    ds.openAttribute("HSSOA_size").read(H5::PredType::NATIVE_UINT, &s.n);
    ds.openAttribute("nPtsArrays").read(H5::PredType::NATIVE_INT, &s.lines);
    s.elem = ds.getDataType().getSize();
    s.cohesion = H5Lexists(f.getId(), "cohesion", H5P_DEFAULT) > 0;
    return s;
}

}  // namespace

extern "C" {

const char *ice_io_last_error(void) { return g_err.c_str(); }

int ice_snapshot_save(const char *path, const float *soa, size_t pitch, uint32_t n, int n_lines, int simulation_step,
                      double simulation_time, int precision) {
    if (!path || (n && !soa) || n_lines < kRefLines || n_lines > kRefLines + 1 || pitch < n ||
        (precision != 32 && precision != 64))
        return fail(ICE_IO_E_ARG, "ice_snapshot_save: bad arguments");
    try {
        H5::Exception::dontPrint();
        H5::H5File file(path, H5F_ACC_TRUNC);
        const hsize_t dims[2] = {(hsize_t)kRefLines, (hsize_t)n};
        H5::DataSpace space(2, dims);
        const hsize_t chunk[2] = {(hsize_t)kRefLines, (hsize_t)std::min<uint32_t>(256u * 1024u, std::max(n, 1u))};
        const H5::PredType &ftype = precision == 32 ? H5::PredType::NATIVE_FLOAT : H5::PredType::NATIVE_DOUBLE;
        H5::DataSet ds = file.createDataSet("pts_data", ftype, space, chunked(2, chunk));
        if (n) {
            if (precision == 32) {   // straight from the SoA lines: a hyperslab of [n_lines][pitch]
                const hsize_t mdims[2] = {(hsize_t)n_lines, (hsize_t)pitch};
                H5::DataSpace mem(2, mdims);
                const hsize_t off[2] = {0, 0}, cnt[2] = {(hsize_t)kRefLines, (hsize_t)n};
                mem.selectHyperslab(H5S_SELECT_SET, cnt, off);
                ds.write(soa, H5::PredType::NATIVE_FLOAT, mem, space);
            } else {
                std::vector<double> buf((size_t)kRefLines * n);
                for (int l = 0; l < kRefLines; l++)
                    for (uint32_t i = 0; i < n; i++) {
                        const float v = soa[(size_t)l * pitch + i];
                        buf[(size_t)l * n + i] = l < kIntLines ? int_in_double(float_bits(v)) : (double)v;
                    }
                ds.write(buf.data(), H5::PredType::NATIVE_DOUBLE);
            }
        }
        write_scalar_attr(ds, "SimulationStep", H5::PredType::NATIVE_INT, simulation_step);
        write_scalar_attr(ds, "SimulationTime", H5::PredType::NATIVE_DOUBLE, simulation_time);
        write_scalar_attr(ds, "HSSOA_size", H5::PredType::NATIVE_UINT, n);
        write_scalar_attr(ds, "nPtsArrays", H5::PredType::NATIVE_INT, kRefLines);
        if (n_lines == kRefLines + 1) {
            const hsize_t cd[1] = {(hsize_t)n}, cc[1] = {chunk[1]};
            H5::DataSpace cs(1, cd);
            H5::DataSet dc = file.createDataSet("cohesion", H5::PredType::NATIVE_FLOAT, cs, chunked(1, cc));
            if (n) dc.write(soa + (size_t)kRefLines * pitch, H5::PredType::NATIVE_FLOAT);
        }
    } catch (const H5::Exception &e) {
        return fail(ICE_IO_E_FILE, std::string("ice_snapshot_save: ") + e.getDetailMsg());
    }
    return 0;
}

int ice_snapshot_info(const char *path, uint32_t *n, int *simulation_step, double *simulation_time, int *precision,
                      int *has_cohesion) {
    if (!path) return fail(ICE_IO_E_ARG, "ice_snapshot_info: no path");
    try {
        H5::Exception::dontPrint();
        H5::H5File file(path, H5F_ACC_RDONLY);
        H5::DataSet ds = file.openDataSet("pts_data");
        const Snap s = read_header(file, ds);
        if (n) *n = s.n;
        if (simulation_step) *simulation_step = s.step;
        if (simulation_time) *simulation_time = s.time;
        if (precision) *precision = (int)(8 * s.elem);
        if (has_cohesion) *has_cohesion = s.cohesion ? 1 : 0;
    } catch (const H5::Exception &e) {
        return fail(ICE_IO_E_FILE, std::string("ice_snapshot_info: ") + e.getDetailMsg());
    }
    return 0;
}

int ice_snapshot_load(const char *path, float *soa, size_t pitch, uint32_t capacity, int n_lines, uint32_t *n,
                      int *simulation_step, double *simulation_time) {
    if (!path || !soa || n_lines < kRefLines || n_lines > kRefLines + 1 || pitch < capacity)
        return fail(ICE_IO_E_ARG, "ice_snapshot_load: bad arguments");
    try {
        H5::Exception::dontPrint();
        H5::H5File file(path, H5F_ACC_RDONLY);
        H5::DataSet ds = file.openDataSet("pts_data");
        const Snap s = read_header(file, ds);
        H5::DataSpace fsp = ds.getSpace();
        hsize_t dims[2] = {0, 0};
        if (fsp.getSimpleExtentNdims() != 2) return fail(ICE_IO_E_FORMAT, "pts_data is not two-dimensional");
        fsp.getSimpleExtentDims(dims);
        if (dims[0] != (hsize_t)kRefLines || dims[1] != s.n || s.lines != kRefLines)   // snapshotmanager.cpp:324-325
            return fail(ICE_IO_E_FORMAT, "array size mismatch");
        if (s.n > capacity) return fail(ICE_IO_E_ARG, "the snapshot holds more points than the capacity");
        if (s.elem != 4 && s.elem != 8) return fail(ICE_IO_E_FORMAT, "pts_data is neither float nor double");
        if (s.n) {
            if (s.elem == 4) {
                const hsize_t mdims[2] = {(hsize_t)n_lines, (hsize_t)pitch};
                H5::DataSpace mem(2, mdims);
                const hsize_t off[2] = {0, 0}, cnt[2] = {(hsize_t)kRefLines, (hsize_t)s.n};
                mem.selectHyperslab(H5S_SELECT_SET, cnt, off);
                ds.read(soa, H5::PredType::NATIVE_FLOAT, mem, fsp);
            } else {
                std::vector<double> buf((size_t)kRefLines * s.n);
                ds.read(buf.data(), H5::PredType::NATIVE_DOUBLE);
                for (int l = 0; l < kRefLines; l++)
                    for (uint32_t i = 0; i < s.n; i++) {
                        const double v = buf[(size_t)l * s.n + i];
                        soa[(size_t)l * pitch + i] = l < kIntLines ? bits_float(int_from_double(v)) : (float)v;
                    }
            }
        }
        if (n_lines == kRefLines + 1) {
            float *coh = soa + (size_t)kRefLines * pitch;
            if (s.cohesion && s.n) {
                H5::DataSet dc = file.openDataSet("cohesion");
                hsize_t cd[1] = {0};
                dc.getSpace().getSimpleExtentDims(cd);
                if (cd[0] != s.n) return fail(ICE_IO_E_FORMAT, "cohesion size mismatch");
                dc.read(coh, H5::PredType::NATIVE_FLOAT);
            } else {
                std::fill(coh, coh + s.n, 0.f);
            }
        }
        if (n) *n = s.n;
        if (simulation_step) *simulation_step = s.step;
        if (simulation_time) *simulation_time = s.time;
    } catch (const H5::Exception &e) {
        return fail(ICE_IO_E_FILE, std::string("ice_snapshot_load: ") + e.getDetailMsg());
    }
    return 0;
}

int ice_read_static_flow(const char *path, int width, int height, int ox, int oy, int gx, int gy, float *current) {
    if (!path || !current || width <= 0 || height <= 0 || gx <= 0 || gy <= 0 || ox < 0 || oy < 0 ||
        ox + gx > width || oy + gy > height)
        return fail(ICE_IO_E_ARG, "ice_read_static_flow: bad arguments (the grid must lie inside the image)");
    try {
        H5::Exception::dontPrint();
        H5::H5File file(path, H5F_ACC_RDONLY);
        H5::DataSet ds = file.openDataSet("velocities");
        H5::DataSpace sp = ds.getSpace();
        if (sp.getSimpleExtentNdims() != 3) return fail(ICE_IO_E_FORMAT, "Incorrect dataset rank");
        hsize_t d[3];
        sp.getSimpleExtentDims(d);
        if (d[0] != (hsize_t)height || d[1] != (hsize_t)width || d[2] != 2)
            return fail(ICE_IO_E_FORMAT, "ReadStaticFlowData: dimension mismatch");
        std::vector<float> v((size_t)width * height * 2);
        ds.read(v.data(), H5::PredType::NATIVE_FLOAT);
        const size_t N = (size_t)gx * gy;
        for (int i = 0; i < gx; i++)
            for (int j = 0; j < gy; j++) {
                const size_t idx = (size_t)(i + ox) + (size_t)(j + oy) * width;
                const size_t idx2 = (size_t)j + (size_t)i * gy;
                current[idx2] = v[2 * idx];
                current[N + idx2] = v[2 * idx + 1];
            }
    } catch (const H5::Exception &e) {
        return fail(ICE_IO_E_FILE, std::string("ice_read_static_flow: ") + e.getDetailMsg());
    }
    return 0;
}

int ice_write_static_flow(const char *path, int width, int height, const float *velocities) {
    if (!path || !velocities || width <= 0 || height <= 0) return fail(ICE_IO_E_ARG, "ice_write_static_flow: bad arguments");
    try {
        H5::Exception::dontPrint();
        H5::H5File file(path, H5F_ACC_TRUNC);
        const hsize_t d[3] = {(hsize_t)height, (hsize_t)width, 2};
        H5::DataSpace sp(3, d);
        H5::DataSet ds = file.createDataSet("velocities", H5::PredType::NATIVE_FLOAT, sp);
        ds.write(velocities, H5::PredType::NATIVE_FLOAT);
    } catch (const H5::Exception &e) {
        return fail(ICE_IO_E_FILE, std::string("ice_write_static_flow: ") + e.getDetailMsg());
    }
    return 0;
}

}  // extern "C"
