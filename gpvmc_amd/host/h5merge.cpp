// The merge of output files into samples (tools/concatenate.cpp): record grouping and the
// HDF5 copy.  Host code only; no device work.
#include "gpvmc_host.hpp"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <algorithm>
#include <string>
#include <vector>

namespace gpvmc {

RecordBunches BunchRecords(int nsamp, const std::vector<int> &statistics)
{
    if (nsamp <= 0) throw Error("the number of samples must be positive");
    if ((size_t)nsamp > statistics.size())
        throw Error("more samples (" + std::to_string(nsamp) + ") than records (" +
                    std::to_string(statistics.size()) + ")");
    // indices in increasing statistics; the same std::sort call as concatenate.cpp:112-115,
    // so that records of equal statistics come out in the same order
    std::vector<int> order(statistics.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return statistics[a] < statistics[b]; });
    RecordBunches b;
    b.members.assign(nsamp, {});
    b.statistics.assign(nsamp, 0);
    size_t lo = 0, hi = order.size();  // the records not yet placed: order[lo .. hi)
    int n = 0;
    while (hi - lo > (size_t)(2 * nsamp + 1)) {
        const int small = order[lo++], large = order[--hi];
        b.members[n].push_back(small);
        b.members[n].push_back(large);
        b.statistics[n] += statistics[small] + statistics[large];
        if (++n == nsamp) n = 0;
    }
    while (hi > lo) {
        const int r = order[--hi];
        const int s = (int)(std::min_element(b.statistics.begin(), b.statistics.end()) - b.statistics.begin());
        b.members[s].push_back(r);
        b.statistics[s] += statistics[r];
    }
    for (int s = 0; s < nsamp; ++s)
        if (b.members[s].empty()) throw Error("sample " + std::to_string(s) + " received no record");
    return b;
}

namespace {

// an HDF5 id closed at scope exit
struct H5Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    H5Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H5Id()
    {
        if (id >= 0 && close) close(id);
    }
    H5Id(const H5Id &) = delete;
    H5Id &operator=(const H5Id &) = delete;
};

struct Record {
    size_t file;
    std::string path;
};

herr_t copy_root_attribute(hid_t loc, const char *name, const H5A_info_t *, void *op)
{
    const hid_t out = *static_cast<hid_t *>(op);
    H5Id a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    if (a.id < 0) return -1;
    H5Id type(H5Aget_type(a.id), H5Tclose);
    H5Id space(H5Aget_space(a.id), H5Sclose);
    if (type.id < 0 || space.id < 0) return -1;
    std::vector<char> buf(std::max<hsize_t>(H5Aget_storage_size(a.id), 1));
    if (H5Aread(a.id, type.id, buf.data()) < 0) return -1;
    H5Id o(H5Acreate2(out, name, type.id, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (o.id < 0 || H5Awrite(o.id, type.id, buf.data()) < 0) return -1;
    return 0;
}

}  // namespace

void Concatenate(const std::vector<std::string> &in, const std::string &out, int nsamp)
{
    if (in.empty()) throw Error("no input file");
    if (nsamp < 0) throw Error("the number of samples must not be negative");
    std::vector<hid_t> files;
    struct Closer {
        std::vector<hid_t> &f;
        ~Closer()
        {
            for (hid_t x : f) H5Fclose(x);
        }
    } closer{files};
    for (const std::string &name : in) {
        const hid_t f = H5Fopen(name.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
        if (f < 0) throw Error("could not open file \"" + name + "\"");
        files.push_back(f);
    }
    // the records and their statistics (concatenate.cpp:135-160: the attribute is read as a
    // double and kept as an int)
    std::vector<Record> records;
    std::vector<int> stats;
    for (size_t f = 0; f < files.size(); ++f)
        for (int r = 1;; ++r) {
            const std::string group = "/rank-" + std::to_string(r);
            if (H5Lexists(files[f], group.c_str(), H5P_DEFAULT) <= 0) break;
            for (int k = 0;; ++k) {
                const std::string path = group + "/data-" + std::to_string(k);
                if (H5Lexists(files[f], path.c_str(), H5P_DEFAULT) <= 0) break;
                double s = 0;
                if (H5LTget_attribute_double(files[f], path.c_str(), "statistics", &s) < 0)
                    throw Error(in[f] + ":" + path + " has no statistics attribute");
                records.push_back({f, path});
                stats.push_back((int)s);
            }
        }
    if (records.empty()) throw Error("no /rank-<r>/data-<k> record (r >= 1) in the input files");
    const RecordBunches b = BunchRecords(nsamp ? nsamp : (int)records.size(), stats);

    const hid_t fo = H5Fcreate(out.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
    if (fo < 0) throw Error("could not create file \"" + out + "\"");
    H5Id fout(fo, H5Fclose);
    hid_t fo_id = fo;
    if (H5Aiterate2(files[0], H5_INDEX_NAME, H5_ITER_NATIVE, nullptr, copy_root_attribute, &fo_id) < 0)
        throw Error("could not copy the root attributes of \"" + in[0] + "\"");
    H5Id group(H5Gcreate2(fo, "/rank-1", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (group.id < 0) throw Error("could not create /rank-1");
    for (size_t s = 0; s < b.members.size(); ++s) {
        const std::vector<int> &m = b.members[s];
        const Record &first = records[m[0]];
        H5Id d0(H5Dopen2(files[first.file], first.path.c_str(), H5P_DEFAULT), H5Dclose);
        if (d0.id < 0) throw Error("could not open " + first.path);
        H5Id space(H5Dget_space(d0.id), H5Sclose);
        H5Id type(H5Dget_type(d0.id), H5Tclose);
        const hssize_t n = H5Sget_simple_extent_npoints(space.id);
        if (n < 0) throw Error("could not read the shape of " + first.path);
        std::vector<double> sum((size_t)n), buf((size_t)n);
        if (H5Dread(d0.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, sum.data()) < 0)
            throw Error("could not read " + first.path);
        for (size_t i = 1; i < m.size(); ++i) {
            const Record &rec = records[m[i]];
            H5Id d(H5Dopen2(files[rec.file], rec.path.c_str(), H5P_DEFAULT), H5Dclose);
            if (d.id < 0) throw Error("could not open " + rec.path);
            H5Id sp(H5Dget_space(d.id), H5Sclose);
            if (H5Sget_simple_extent_npoints(sp.id) != n)
                throw Error(in[rec.file] + ":" + rec.path + " differs in size from " + in[first.file] + ":" + first.path);
            if (H5Dread(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
                throw Error("could not read " + rec.path);
            for (size_t j = 0; j < sum.size(); ++j) sum[j] += buf[j];
        }
        for (double &v : sum) v /= (double)m.size();
        const std::string path = "/rank-1/data-" + std::to_string(s);
        H5Id dout(H5Dcreate2(fo, path.c_str(), type.id, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        if (dout.id < 0 || H5Dwrite(dout.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, sum.data()) < 0)
            throw Error("could not write " + path);
        if (H5LTset_attribute_int(fo, path.c_str(), "statistics", &b.statistics[s], 1) < 0)
            throw Error("could not write the statistics of " + path);
    }
}

}  // namespace gpvmc
