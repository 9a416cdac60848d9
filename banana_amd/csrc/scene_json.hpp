// scene_json.hpp -- the reference's scene file format (JSON) mapped onto bnn_flip_params and a
// bnn_scene_spec. Host code only.
//
// Keys follow ParticleSolver::loadScene (ParticleSolver.cpp:30-168): GlobalParameters
// (ParticleSolverData.cpp:33-78), SimulationParameters (:148-211) with SimulationDomainBox,
// PIC_Parameters (PIC_Data.cpp:33-38) and ParticleGenerators (ParticleGenerator.cpp:43-55,
// SimulationObject.cpp:38-66, GeometryObjects.cpp:142-170). Keys are case-sensitive as in
// JSONHelpers::readValue (a mis-cased key keeps its default); unknown keys are reported as warnings.
#pragma once

#include <cctype>
#include <cstring>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bnn_flip.h"

namespace bnn {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool                         b = false;
    double                       num = 0;
    std::string                  str;
    std::vector<Value>           arr;
    std::map<std::string, Value> obj;
    const Value* find(const std::string& k) const {
        if(kind != Object) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

class Parser {
public:
    explicit Parser(const std::string& s) : s_(s) {}
    Value parse() {
        Value v = value();
        ws();
        if(i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t             i_ = 0;
    [[noreturn]] void  fail(const char* m) { throw std::runtime_error(std::string("scene JSON: ") + m + " at offset " + std::to_string(i_)); }
    void ws() {
        for(;;) {
            while(i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
            // the reference's scene files carry // comments; accept them
            if(i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while(i_ < s_.size() && s_[i_] != '\n') ++i_;
                continue;
            }
            return;
        }
    }
    Value value() {
        ws();
        if(i_ >= s_.size()) fail("unexpected end");
        const char c = s_[i_];
        if(c == '{') return object();
        if(c == '[') return array();
        if(c == '"') {
            Value v;
            v.kind = Value::String;
            v.str  = string();
            return v;
        }
        if(s_.compare(i_, 4, "true") == 0) {
            i_ += 4;
            Value v;
            v.kind = Value::Bool;
            v.b    = true;
            return v;
        }
        if(s_.compare(i_, 5, "false") == 0) {
            i_ += 5;
            Value v;
            v.kind = Value::Bool;
            return v;
        }
        if(s_.compare(i_, 4, "null") == 0) {
            i_ += 4;
            return Value();
        }
        return number();
    }
    std::string string() {
        if(s_[i_] != '"') fail("expected string");
        ++i_;
        std::string out;
        while(i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if(c == '\\') {
                if(i_ >= s_.size()) fail("bad escape");
                char e = s_[i_++];
                switch(e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u':
                        if(i_ + 4 > s_.size()) fail("bad unicode escape");
                        out += '?';
                        i_ += 4;
                        break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if(i_ >= s_.size()) fail("unterminated string");
        ++i_;
        return out;
    }
    Value number() {
        const char* b   = s_.c_str() + i_;
        char*       end = nullptr;
        double      d   = std::strtod(b, &end);
        if(end == b) fail("unexpected character");
        i_ += (size_t)(end - b);
        Value v;
        v.kind = Value::Number;
        v.num  = d;
        return v;
    }
    Value array() {
        Value v;
        v.kind = Value::Array;
        ++i_;
        ws();
        if(i_ < s_.size() && s_[i_] == ']') {
            ++i_;
            return v;
        }
        for(;;) {
            v.arr.push_back(value());
            ws();
            if(i_ >= s_.size()) fail("unterminated array");
            if(s_[i_] == ',') {
                ++i_;
                continue;
            }
            if(s_[i_] == ']') {
                ++i_;
                return v;
            }
            fail("expected , or ]");
        }
    }
    Value object() {
        Value v;
        v.kind = Value::Object;
        ++i_;
        ws();
        if(i_ < s_.size() && s_[i_] == '}') {
            ++i_;
            return v;
        }
        for(;;) {
            ws();
            std::string k = string();
            ws();
            if(i_ >= s_.size() || s_[i_] != ':') fail("expected :");
            ++i_;
            v.obj[k] = value();
            ws();
            if(i_ >= s_.size()) fail("unterminated object");
            if(s_[i_] == ',') {
                ++i_;
                continue;
            }
            if(s_[i_] == '}') {
                ++i_;
                return v;
            }
            fail("expected , or }");
        }
    }
};

}  // namespace json

struct SceneLoadError : std::runtime_error {
    bnn_status code;
    SceneLoadError(bnn_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Reads a reference scene (already parsed) into params/spec; returns warnings (one per line).
// AdditionalBoundaryObjects (ParticleSolver.cpp:242-258, GeometryObjects.cpp:143-201): box or
// sphere geometry, Translation / RotationAxisAngle (degrees) / Scale, and an Animation with key
// frames. CSGOperator and UniqueName are accepted and not used (the reference merges every static
// object with a union whatever its operator).
inline void boundaries_from_json(const json::Value& root, std::vector<bnn_boundary_object>& out) {
    using json::Value;
    out.clear();
    const Value* ab = root.find("AdditionalBoundaryObjects");
    if(!ab) return;
    if(ab->kind != Value::Array) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: AdditionalBoundaryObjects must be an array");
    auto vec = [](const Value& o, const char* k, float* dst, int n) {
        const Value* v = o.find(k);
        if(!v) return false;
        if(v->kind != Value::Array || (int)v->arr.size() != n)
            throw SceneLoadError(BNN_ERR_INVALID_ARG, std::string("scene JSON: boundary key '") + k + "' has the wrong shape");
        for(int d = 0; d < n; ++d) dst[d] = (float)v->arr[d].num;
        return true;
    };
    auto number = [](const Value& o, const char* k, float& dst) {
        const Value* v = o.find(k);
        if(!v) return false;
        if(v->kind != Value::Number) throw SceneLoadError(BNN_ERR_INVALID_ARG, std::string("scene JSON: boundary key '") + k + "' must be a number");
        dst = (float)v->num;
        return true;
    };
    auto flag = [](const Value& o, const char* k, int32_t& dst) {
        const Value* v = o.find(k);
        if(!v) return false;
        if(v->kind == Value::Bool) dst = v->b ? 1 : 0;
        else if(v->kind == Value::Number) dst = v->num != 0 ? 1 : 0;
        else throw SceneLoadError(BNN_ERR_INVALID_ARG, std::string("scene JSON: boundary key '") + k + "' must be a boolean");
        return true;
    };
    for(const Value& jo : ab->arr) {
        if(jo.kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: a boundary object must be an object");
        if(out.size() >= BNN_MAX_BOUNDARY_OBJECTS) throw SceneLoadError(BNN_ERR_UNSUPPORTED, "scene JSON: too many boundary objects");
        bnn_boundary_object o;
        std::memset(&o, 0, sizeof(o));
        for(int d = 0; d < 3; ++d) {
            o.box_min[d] = -1.f;
            o.box_max[d] = 1.f;
        }
        o.rotation_axis_angle[0] = o.rotation_axis_angle[1] = o.rotation_axis_angle[2] = 1.f;
        o.scale             = 1.f;
        o.cubic_translation = o.cubic_rotation = o.cubic_scale = 1;
        const Value* ty = jo.find("GeometryType");
        if(!ty || ty->kind != Value::String) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: boundary object without GeometryType");
        const std::string& g = ty->str;
        if(g == "Box" || g == "box" || g == "BOX") o.geometry = BNN_GEOMETRY_BOX;
        else if(g == "Sphere" || g == "sphere" || g == "SPHERE") o.geometry = BNN_GEOMETRY_SPHERE;
        else throw SceneLoadError(BNN_ERR_UNSUPPORTED, "scene JSON: boundary geometry '" + g + "' is not provided (Box, Sphere)");
        for(const char* k : {"RotationEulerAngles", "RotationEulerAngle", "MeshFile"})
            if(jo.find(k)) throw SceneLoadError(BNN_ERR_UNSUPPORTED, std::string("scene JSON: boundary key '") + k + "' is not provided");
        vec(jo, "BoxMin", o.box_min, 3);
        vec(jo, "BoxMax", o.box_max, 3);
        o.has_translation = vec(jo, "Translation", o.translation, 3) ? 1 : 0;
        o.has_rotation    = vec(jo, "RotationAxisAngle", o.rotation_axis_angle, 4) ? 1 : 0;
        o.has_scale       = number(jo, "Scale", o.scale) ? 1 : 0;
        if(const Value* an = jo.find("Animation")) {
            if(an->kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: Animation must be an object");
            flag(*an, "CubicInterpolationTranslation", o.cubic_translation);
            flag(*an, "CubicInterpolationRotation", o.cubic_rotation);
            flag(*an, "CubicInterpolationScale", o.cubic_scale);
            if(flag(*an, "Periodic", o.periodic)) {
                float sf = 0;
                if(number(*an, "StartFrame", sf)) o.start_frame = (uint32_t)sf;
            }
            const Value* kf = an->find("KeyFrames");
            if(!kf || kf->kind != Value::Array) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: Animation without KeyFrames");
            if(kf->arr.size() > BNN_MAX_KEYFRAMES) throw SceneLoadError(BNN_ERR_UNSUPPORTED, "scene JSON: too many key frames");
            for(const Value& jk : kf->arr) {
                if(jk.kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: a key frame must be an object");
                for(const char* k : {"RotationEulerAngles", "RotationEulerAngle"})
                    if(jk.find(k)) throw SceneLoadError(BNN_ERR_UNSUPPORTED, std::string("scene JSON: key frame key '") + k + "' is not provided");
                bnn_keyframe& k = o.keyframes[o.n_keyframes++];
                float         fr = 0;
                if(!number(jk, "Frame", fr)) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: a key frame without Frame");
                k.frame = (uint32_t)fr;
                vec(jk, "Translation", k.translation, 3);
                k.has_rotation = vec(jk, "RotationAxisAngle", k.rotation_axis_angle, 4) ? 1 : 0;
                k.scale        = 1.f;
                k.has_scale    = number(jk, "Scale", k.scale) ? 1 : 0;
            }
        }
        out.push_back(o);
    }
}

inline std::string scene_from_json(const json::Value& root, bnn_flip_params& p, bnn_scene_spec& sp) {
    using json::Value;
    std::string warn;
    auto        note = [&](const std::string& w) { warn += w + "\n"; };
    if(root.kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: the top level must be an object");
    const Value* gp = root.find("GlobalParameters");
    const Value* sm = root.find("SimulationParameters");
    if(!gp || gp->kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: GlobalParameters is required");
    if(!sm || sm->kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: SimulationParameters is required");

    // ParticleSolverFactory::createSolverFromJSon (ParticleSolverFactory.cpp:59-75): the 3-D PIC-family
    // solvers whose step this library provides, by their factory names (FLIP_Solver.h:37,
    // APIC_Solver.h, AFLIP_Solver.h)
    if(const Value* s = gp->find("Solver")) {
        const std::string name = s->kind == Value::String ? s->str : std::string("?");
        if(name == "FLIP_3DSolver") p.transfer_scheme = BNN_TRANSFER_FLIP;
        else if(name == "APIC_3DSolver") p.transfer_scheme = BNN_TRANSFER_APIC;
        else if(name == "AFLIP_3DSolver") p.transfer_scheme = BNN_TRANSFER_AFLIP;
        else
            throw SceneLoadError(BNN_ERR_UNSUPPORTED, "scene JSON: solver '" + name +
                                                          "' is not provided (this library is FLIP_3DSolver, APIC_3DSolver, AFLIP_3DSolver)");
    }
    auto num = [&](const Value& o, const char* k, auto& dst) -> bool {
        const Value* v = o.find(k);
        if(!v) return false;
        if(v->kind == Value::Number) {
            dst = (std::decay_t<decltype(dst)>)v->num;
            return true;
        }
        if(v->kind == Value::Bool) {
            dst = (std::decay_t<decltype(dst)>)(v->b ? 1 : 0);
            return true;
        }
        note(std::string("key '") + k + "' has the wrong type; default kept");
        return false;
    };
    auto vec3 = [&](const Value& o, const char* k, float* dst) -> bool {
        const Value* v = o.find(k);
        if(!v) return false;
        if(v->kind != Value::Array || v->arr.size() != 3) {
            note(std::string("key '") + k + "' is not a 3-vector; default kept");
            return false;
        }
        for(int d = 0; d < 3; ++d) dst[d] = (float)v->arr[d].num;
        return true;
    };

    // GlobalParameters (ParticleSolverData.cpp:33-78)
    static const std::set<std::string> gKnown = {"Solver", "NThreads", "FrameDuration", "StartFrame", "FinalFrame", "DataPath",
                                                 "MemoryStateDataFolder", "FrameDataFolder", "LoadMemoryState", "SaveMemoryState",
                                                 "SaveFrameData", "SaveSubstepData", "FramePerState", "OptionalSavingData", "LogLevel",
                                                 "PrintLogToConsole", "PrintLogToFile", "ApplyGravity", "EnableSortParticle",
                                                 "SortFrequency"};
    for(const auto& kv : gp->obj)
        if(!gKnown.count(kv.first)) note("GlobalParameters: unknown key '" + kv.first + "' ignored");
    num(*gp, "FrameDuration", p.frame_duration);
    num(*gp, "StartFrame", p.start_frame);
    num(*gp, "FinalFrame", p.final_frame);
    num(*gp, "ApplyGravity", p.apply_gravity);
    num(*gp, "EnableSortParticle", p.enable_sort_particle);
    num(*gp, "SortFrequency", p.sort_frequency);

    // SimulationParameters (ParticleSolverData.cpp:148-211, PIC_Data.cpp:33-38)
    static const std::set<std::string> sKnown = {
        "SimulationDomainBox", "MinTimestep", "MaxTimestep", "CFLFactor", "ParticleRadius", "CellSize", "RatioCellSizePRadius",
        "NExpandCells", "MaxNParticles", "OverlapThresholdRatio", "CollisionThresholdRatio", "MaterialDensity", "AdvectionSteps",
        "CorrectPosition", "RepulsiveForceStiffness", "CGRelativeTolerance", "MaxCGIteration", "PIC_FLIP_ratio",
        "ReflectVelocityAtBoundary", "BoundaryReflectionMultiplier", "GravityType", "GravityDirection", "GravityCenter",
        "ExitIfPressureProjectionFailed", "BoundarySDFAtInit"};
    for(const auto& kv : sm->obj)
        if(!sKnown.count(kv.first)) note("SimulationParameters: unknown key '" + kv.first + "' ignored (keys are case-sensitive)");
    num(*sm, "MinTimestep", p.min_timestep);
    num(*sm, "MaxTimestep", p.max_timestep);
    num(*sm, "CFLFactor", p.cfl_factor);
    num(*sm, "CellSize", p.cell_size);
    num(*sm, "RatioCellSizePRadius", p.ratio_cell_size_pradius);
    num(*sm, "NExpandCells", p.n_expand_cells);
    num(*sm, "AdvectionSteps", p.advection_steps);
    num(*sm, "CorrectPosition", p.correct_position);
    num(*sm, "RepulsiveForceStiffness", p.repulsive_force_stiffness);
    num(*sm, "CGRelativeTolerance", p.cg_relative_tolerance);
    num(*sm, "MaxCGIteration", p.max_cg_iteration);
    num(*sm, "PIC_FLIP_ratio", p.pic_flip_ratio);
    num(*sm, "ReflectVelocityAtBoundary", p.reflect_velocity_at_boundary);
    num(*sm, "BoundaryReflectionMultiplier", p.boundary_reflection_multiplier);
    num(*sm, "ExitIfPressureProjectionFailed", p.exit_if_pressure_projection_failed);
    num(*sm, "BoundarySDFAtInit", p.boundary_sdf_at_init);
    if(const Value* gt = sm->find("GravityType")) {
        if(gt->kind != Value::String) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: GravityType must be a string");
        static const char* names[] = {"Earth", "Directional", "ToCenter", "FromCenter"};
        int                t       = -1;
        for(int i = 0; i < 4; ++i)
            if(gt->str == names[i]) t = i;
        if(t < 0) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: unknown GravityType '" + gt->str + "'");
        p.gravity_type = t;
    }
    vec3(*sm, "GravityDirection", p.gravity_direction);
    vec3(*sm, "GravityCenter", p.gravity_center);

    // domain box (ParticleSolver.cpp:70-109): BoxMin/BoxMax, uniform Scale, Translation
    const Value* box = sm->find("SimulationDomainBox");
    if(!box || box->kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: SimulationDomainBox is required");
    float bmin[3] = {-1, -1, -1}, bmax[3] = {1, 1, 1}, t[3] = {0, 0, 0}, s = 1.f;
    vec3(*box, "BoxMin", bmin);
    vec3(*box, "BoxMax", bmax);
    vec3(*box, "Translation", t);
    num(*box, "Scale", s);
    for(int d = 0; d < 3; ++d) {
        p.box_min[d] = s * bmin[d] + t[d];
        p.box_max[d] = s * bmax[d] + t[d];
    }

    if(const Value* ab = root.find("AdditionalBoundaryObjects")) {  // parsed by boundaries_from_json
        std::vector<bnn_boundary_object> objs;
        boundaries_from_json(root, objs);
    }
    if(const Value* pr = root.find("ParticleRemovers"))
        if(pr->kind == Value::Array && !pr->arr.empty()) note("ParticleRemovers are not provided; ignored");

    // ParticleGenerators: Box (BoxMin/BoxMax) and Sphere (unit sphere scaled and translated)
    const Value* gens = root.find("ParticleGenerators");
    if(!gens || gens->kind != Value::Array || gens->arr.empty())
        throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: ParticleGenerators is required");
    std::memset(&sp, 0, sizeof(sp));
    sp.ppa    = 2;
    sp.jitter = 1.0f;
    for(const Value& gv : gens->arr) {
        if(gv.kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: a generator must be an object");
        const Value* ty = gv.find("GeometryType");
        if(!ty || ty->kind != Value::String) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: generator without GeometryType");
        for(const char* k : {"RotationEulerAngles", "RotationEulerAngle", "RotationAxisAngle", "Animation"})
            if(gv.find(k)) throw SceneLoadError(BNN_ERR_UNSUPPORTED, std::string("scene JSON: generator key '") + k + "' is not provided");
        float gt[3] = {0, 0, 0}, gs = 1.f, v0[3] = {0, 0, 0};
        vec3(gv, "Translation", gt);
        num(gv, "Scale", gs);
        vec3(gv, "InitialVelocity", v0);
        const std::string& g = ty->str;
        if(g == "Box" || g == "box" || g == "BOX") {
            if(sp.n_boxes >= 8) throw SceneLoadError(BNN_ERR_UNSUPPORTED, "scene JSON: at most 8 box generators");
            float lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};
            vec3(gv, "BoxMin", lo);
            vec3(gv, "BoxMax", hi);
            const int i = sp.n_boxes++;
            for(int d = 0; d < 3; ++d) {
                sp.box_min[i][d] = gs * lo[d] + gt[d];
                sp.box_max[i][d] = gs * hi[d] + gt[d];
                sp.box_vel[i][d] = v0[d];
            }
        } else if(g == "Sphere" || g == "sphere" || g == "SPHERE") {
            if(sp.n_spheres >= 8) throw SceneLoadError(BNN_ERR_UNSUPPORTED, "scene JSON: at most 8 sphere generators");
            const int i          = sp.n_spheres++;
            sp.sphere_radius[i] = gs;
            for(int d = 0; d < 3; ++d) {
                sp.sphere_center[i][d] = gt[d];
                sp.sphere_vel[i][d]    = v0[d];
            }
        } else {
            throw SceneLoadError(BNN_ERR_UNSUPPORTED, "scene JSON: generator geometry '" + g + "' is not provided (Box, Sphere)");
        }
    }
    return warn;
}

// The file IO keys of GlobalParameters (ParticleSolverData.cpp:48-56): data path, folders, load /
// save switches, frames per state, and the OptionalSavingData list. Wrongly typed values keep their
// defaults.
inline void io_from_json(const json::Value& root, bnn_io_params& io) {
    using json::Value;
    if(root.kind != Value::Object) throw SceneLoadError(BNN_ERR_INVALID_ARG, "scene JSON: the top level must be an object");
    const Value* gp = root.find("GlobalParameters");
    if(!gp || gp->kind != Value::Object) return;
    auto str = [&](const char* k, char* dst, size_t cap) {
        const Value* v = gp->find(k);
        if(v && v->kind == Value::String && v->str.size() + 1 <= cap) std::memcpy(dst, v->str.c_str(), v->str.size() + 1);
    };
    auto flag = [&](const char* k, int32_t& dst) {
        const Value* v = gp->find(k);
        if(v && v->kind == Value::Bool) dst = v->b ? 1 : 0;
        if(v && v->kind == Value::Number) dst = v->num != 0 ? 1 : 0;
    };
    str("DataPath", io.data_path, sizeof(io.data_path));
    str("MemoryStateDataFolder", io.memory_state_folder, sizeof(io.memory_state_folder));
    str("FrameDataFolder", io.frame_data_folder, sizeof(io.frame_data_folder));
    flag("LoadMemoryState", io.load_memory_state);
    flag("SaveMemoryState", io.save_memory_state);
    flag("SaveFrameData", io.save_frame_data);
    if(const Value* v = gp->find("FramePerState"); v && v->kind == Value::Number && v->num >= 1) io.frame_per_state = (uint32_t)v->num;
    if(const Value* v = gp->find("OptionalSavingData"); v && v->kind == Value::Array)
        for(const Value& e : v->arr) {
            if(e.kind != Value::String) continue;
            if(e.str == "ParticleVelocity") io.save_velocity = 1;
            if(e.str == "ObjectIndex") io.save_object_index = 1;
        }
}

}  // namespace bnn
