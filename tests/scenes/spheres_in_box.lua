-- A scene in the reference's scene-script language, written for this
-- repository's tests: a closed box of Lambertian triangles with a light quad,
-- three spheres with the reference's four BSDF kinds (one layered), a smooth
-- triangle, and a loop that places a row of small spheres. It loads the
-- scene functions through `require "core"` like the reference's samples do;
-- without the reference's lua_core directory on package.path the library's
-- built-in scene functions answer.
package.path = package.path .. ';../lua_core/?.lua'
package.path = package.path .. ';./lua_core/?.lua'

require "core"

local w, h = 48, 36

Camera{
    type = "pinhole",
    position = {0, 0, 3.5},
    look_at = {0, 0, 0},
    fov = 45
}

PixelSampler{ type = "jittered", spp = 9 }

ImageBuffer{
    width = w,
    height = h,
    file_name = "spheres_in_box.ppm",
    viewport_width = w,
    viewport_height = h
}

GlobalSettings{ background_color = {0, 0, 0} }

Integrator{
    type = "path-tracer",
    path_termination = "russian-roulette",
    path_length = 3
}

local none = Emission{0, 0, 0}
local function diffuse(r, g, b)
    return Material{ bsdf = LayeredBSDF{ Lambertian{ kd = {r, g, b} } }, emission = none }
end

local white = diffuse(0.7, 0.7, 0.7)
local red = diffuse(0.7, 0.1, 0.1)
local light = Material{ bsdf = LayeredBSDF{ Lambertian{ kd = {0, 0, 0} } }, emission = Emission{6, 6, 6} }
local copper = Material{
    bsdf = LayeredBSDF{ CookTorrance{ roughness = 0.25, reflectance_at_normal_incidence = {0.95, 0.64, 0.54} } },
    emission = none
}
local mirror = Material{ bsdf = LayeredBSDF{ SmoothConductor{ reflectance_at_normal_incidence = {0.9, 0.9, 0.9} } },
                         emission = none }
local plastic = Material{
    bsdf = LayeredBSDF{ SmoothDielectric{ ior = 1.5 }, Lambertian{ kd = {0.2, 0.3, 0.7}, bsdf_sampler = "uniform" } },
    emission = none
}

local a = 1
local function quad(p, q, r, s, m)
    Triangle{ v1 = p, v2 = q, v3 = r, material = m }
    Triangle{ v1 = p, v2 = r, v3 = s, material = m }
end
quad({-a, -a, -a}, { a, -a, -a}, { a,  a, -a}, {-a,  a, -a}, white)  -- back
quad({-a, -a,  a}, { a, -a,  a}, { a, -a, -a}, {-a, -a, -a}, white)  -- floor
quad({-a,  a, -a}, { a,  a, -a}, { a,  a,  a}, {-a,  a,  a}, white)  -- ceiling
quad({-a, -a,  a}, {-a, -a, -a}, {-a,  a, -a}, {-a,  a,  a}, red)    -- left
quad({ a, -a, -a}, { a, -a,  a}, { a,  a,  a}, { a,  a, -a}, white)  -- right
local l, y = 0.6, a - 0.001
quad({-l, y, -l}, {l, y, -l}, {l, y, l}, {-l, y, l}, light)

Sphere{ center = {-0.45, -0.6, -0.3}, radius = 0.4, material = copper }
Sphere{ center = { 0.45, -0.65, 0.2}, radius = 0.35, material = plastic }
Sphere{ center = { 0.1, 0.2, -0.5}, radius = 0.25, material = mirror }

SmoothTriangle{
    v1 = {-0.9, 0.5, -0.9}, v2 = {-0.5, 0.5, -0.9}, v3 = {-0.7, 0.9, -0.8},
    n1 = {0, 0, 1}, n2 = {0.3, 0, 0.95}, n3 = {0, 0.3, 0.95},
    material = white
}

for i = 1, 5 do
    Sphere{ center = {-0.8 + 0.4 * (i - 1), -0.92, 0.7}, radius = 0.07, material = (i % 2 == 0) and copper or plastic }
end
