import sys, time, json, ctypes as C
sys.path.insert(0, ".")
import bench
bench._package()
from cr_amd import crgpu, scenes
new = crgpu.load_library()
prev = crgpu.load_library("cpu-renderer_amd/lib/libcrgpu_prev.so")
def ctx_for(lib):
    c = crgpu.Context.__new__(crgpu.Context)
    c.lib = lib; c.handle = C.c_void_p()
    assert lib.crg_context_create(0, C.byref(c.handle)) == 0
    return c
for wl in ("c2", "c3"):
    w = scenes.config2() if wl == "c2" else scenes.config3()
    ctxs = {"prev": ctx_for(prev), "new": ctx_for(new)}
    ids = {k: (c.upload_scene(w.scene)[0], c.upload_env(w.env)[0]) for k, c in ctxs.items()}
    opt = w.options()
    films = {}
    for rep in range(2):
        for k, c in ctxs.items():
            sid, eid = ids[k]
            c.render(sid, eid, w.camera, opt)
            c.set_profiling(True); r = c.render(sid, eid, w.camera, opt); kt = c.kernel_times(); c.set_profiling(False)
            t0 = time.perf_counter(); r2 = c.render(sid, eid, w.camera, opt); el = time.perf_counter() - t0
            films[k] = r2.pixels.tobytes()
            print(json.dumps({"wl": wl, "lib": k, "render_s": round(el, 4), "msamples": round(w.camera_samples / el / 1e6, 1),
                              "k_extend": round(kt["k_extend"][1], 2), "k_shadow": round(kt["k_shadow"][1], 2), "k_shade": round(kt["k_shade"][1], 2)}), flush=True)
    print(wl, "films identical:", films["prev"] == films["new"], flush=True)
    for k, c in ctxs.items():
        c.release_scene(ids[k][0]); c.release_env(ids[k][1]); c.close()
