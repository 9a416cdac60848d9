// flip_io.hpp -- the reference's particle file format, "BananaParticleData"
// (Simulation/ParticleTools/ParticleSerialization.cpp:533-632 write, 635-830 read), used for frame
// output (PIC_Solver::saveFrameData, PIC_Solver.cpp:232-262) and memory states
// (PIC_Solver::saveMemoryState / loadMemoryState, PIC_Solver.cpp:158-229).
//
// A file is a text header followed by raw little-endian data:
//   BananaParticleData
//   FixedAttribute <name> <type> <element size> <count> <bytes>      (one line per fixed attribute)
//   ParticleAttribute <name> <type> <element size> <count> <bytes>   (one line per particle attribute)
//   NParticles <n>
//   EndHeader.
// then the fixed attributes' bytes, then the particle attributes' bytes, each group in ascending
// order of attribute name (the reference keeps both groups in a std::map). A "compressed_real"
// attribute of count N holds N float minima, N float maxima, then one uint16 per component:
// (uint16)(65535 * ((v - min) / (max - min))) in float arithmetic (ParticleHelpers.cpp:47-79).
#pragma once

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace bnn {

struct BpdAttr {
    std::string          name, type;
    int                  elemSize = 4;
    uint32_t             count    = 1;
    std::vector<uint8_t> data;
};

struct BpdFile {
    std::vector<BpdAttr> fixed, particle;
    uint64_t             nParticles = 0;

    template<class T> void addFixed(const std::string& name, const char* type, uint32_t count, const T* v) {
        BpdAttr a;
        a.name     = name;
        a.type     = type;
        a.elemSize = (int)sizeof(T);
        a.count    = count;
        a.data.resize(sizeof(T) * count);
        std::memcpy(a.data.data(), v, a.data.size());
        fixed.push_back(std::move(a));
    }
    template<class T> void addParticle(const std::string& name, const char* type, uint32_t count, const T* v) {
        BpdAttr a;
        a.name     = name;
        a.type     = type;
        a.elemSize = (int)sizeof(T);
        a.count    = count;
        a.data.resize(sizeof(T) * count * nParticles);
        if(!a.data.empty()) std::memcpy(a.data.data(), v, a.data.size());
        particle.push_back(std::move(a));
    }
    void addRaw(bool isFixed, const std::string& name, const char* type, int elemSize, uint32_t count, std::vector<uint8_t>&& bytes) {
        BpdAttr a;
        a.name     = name;
        a.type     = type;
        a.elemSize = elemSize;
        a.count    = count;
        a.data     = std::move(bytes);
        (isFixed ? fixed : particle).push_back(std::move(a));
    }
    const BpdAttr* find(const std::string& name) const {
        for(const auto* g : {&fixed, &particle})
            for(const auto& a : *g)
                if(a.name == name) return &a;
        return nullptr;
    }
};

// creates the missing directories of path's parent (DataIO::createOutputFolders)
inline bool bpd_make_parent_dirs(const std::string& path) {
    for(size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1)) {
        const std::string d = path.substr(0, p);
        if(::mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
    return true;
}

inline bool bpd_write(const std::string& path, BpdFile f, std::string& err) {
    auto byName = [](const BpdAttr& a, const BpdAttr& b) { return a.name < b.name; };
    std::sort(f.fixed.begin(), f.fixed.end(), byName);
    std::sort(f.particle.begin(), f.particle.end(), byName);
    if(!bpd_make_parent_dirs(path)) {
        err = "cannot create the folders of " + path;
        return false;
    }
    std::ofstream o(path, std::ios::binary | std::ios::out | std::ios::trunc);
    if(!o.is_open()) {
        err = "cannot write " + path;
        return false;
    }
    o.imbue(std::locale::classic());
    o << "BananaParticleData\n";
    for(const auto& a : f.fixed)
        o << "FixedAttribute " << a.name << " " << a.type << " " << a.elemSize << " " << a.count << " " << a.data.size() << "\n";
    for(const auto& a : f.particle)
        o << "ParticleAttribute " << a.name << " " << a.type << " " << a.elemSize << " " << a.count << " " << a.data.size() << "\n";
    o << "NParticles " << f.nParticles << "\n";
    o << "EndHeader.\n";
    for(const auto* g : {&f.fixed, &f.particle})
        for(const auto& a : *g) o.write(reinterpret_cast<const char*>(a.data.data()), (std::streamsize)a.data.size());
    o.close();
    if(!o) {
        err = "write error on " + path;
        return false;
    }
    return true;
}

inline bool bpd_read(const std::string& path, BpdFile& f, std::string& err) {
    std::ifstream in(path, std::ios::binary);
    if(!in.is_open()) {
        err = "cannot read " + path;
        return false;
    }
    f = BpdFile{};
    std::string line;
    bool        magic = false, end = false;
    while(std::getline(in, line)) {
        std::istringstream ls(line);
        ls.imbue(std::locale::classic());
        std::string tok;
        ls >> tok;
        if(tok == "BananaParticleData") {
            magic = true;
        } else if(tok == "EndHeader.") {
            end = true;
            break;
        } else if(tok == "NParticles") {
            ls >> f.nParticles;
        } else if(tok == "FixedAttribute" || tok == "ParticleAttribute") {
            BpdAttr a;
            size_t  bytes = 0;
            ls >> a.name >> a.type >> a.elemSize >> a.count >> bytes;
            if(!ls) {
                err = "malformed header line in " + path + ": " + line;
                return false;
            }
            a.data.resize(bytes);
            (tok == "FixedAttribute" ? f.fixed : f.particle).push_back(std::move(a));
        } else {
            err = "unknown header line in " + path + ": " + line;
            return false;
        }
    }
    if(!magic || !end) {
        err = path + " is not a BananaParticleData file";
        return false;
    }
    // the data follow the header: fixed attributes, then particle attributes, each by name (the
    // writer's map order; the header lists them in the same order)
    for(auto* g : {&f.fixed, &f.particle})
        for(auto& a : *g) {
            in.read(reinterpret_cast<char*>(a.data.data()), (std::streamsize)a.data.size());
            if(!in) {
                err = path + " is truncated (attribute " + a.name + ")";
                return false;
            }
        }
    return true;
}

// DataIO::getFilePath (DataIO.h:119-124): "<root>/<folder>/<name>.%04d.<ext>"
inline std::string bpd_file_path(const std::string& root, const std::string& folder, int frame) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%04d", frame);
    return root + "/" + folder + "/frame." + buf + ".bnn";
}

}  // namespace bnn
