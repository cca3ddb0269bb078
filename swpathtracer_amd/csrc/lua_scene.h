// lua_scene.h - the evaluator behind spt_lua_* (host only, no HIP).
//
// The reference reads its scenes from Lua 5.3 scripts: lua_core/*.lua defines
// the scene functions (Camera{...}, Material{...}, Triangle{...}, ...) that
// validate their tables, fill defaults and register them under
// _G.__RENDERER_ELEMENTS__ (lua_core/registervars.lua:1-16), and
// lua_bind/lua_bind.cpp reads that registry back. liblua is not a dependency
// here, so this header declares a small interpreter for the language subset
// scene scripts use and the scene it extracts.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "spt.h"

namespace spt_lua {

// Extracted scene: what SwpathtracerApp builds from lua_bind's getters
// (core/swpathtracer_app.cpp:40-262), already in the C ABI's layouts.
struct SceneOut {
    spt_lua_info info{};
    std::vector<float> prims;      // n x 9
    std::vector<uint8_t> kinds;    // n
    std::vector<float> normals;    // n x 9 (zeros: flat)
    std::vector<uint32_t> mat_ids; // n
    std::vector<spt_material> mats;
    std::string registry_json;     // _G.__RENDERER_ELEMENTS__ after the script ran
    std::string printed;           // what the script printed
};

// Runs `script` (relative paths resolve against `workdir`, the reference's
// working directory; NULL or "" = the process's) and extracts the scene.
// Throws std::runtime_error with the reference's message on failure.
void load_scene(const std::string& script, const std::string& workdir, uint32_t flags, SceneOut& out);

// Meshes of an OBJ file as Scene::loadMesh adds them (core/scene.cpp:32-68).
struct ObjMesh {
    std::string name;
    std::vector<float> tris;     // m x 9
    std::vector<float> normals;  // m x 9, zeros where a face has no normals
};
std::vector<ObjMesh> read_obj(const std::string& path);

// glibc random() (the additive feedback generator behind Lua 5.3's
// math.random on POSIX), seeded as srandom(seed).
class GlibcRandom {
   public:
    explicit GlibcRandom(uint32_t seed = 1) { seed_with(seed); }
    void seed_with(uint32_t seed);
    int32_t next();

   private:
    int32_t r_[34];
    int k_ = 0;
};

}  // namespace spt_lua
