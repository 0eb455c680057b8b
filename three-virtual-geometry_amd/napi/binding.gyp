{
  "targets": [{
    "target_name": "vgcut",
    "sources": ["vgcut_napi.cc"],
    "include_dirs": ["../../include"],
    "libraries": ["-L<(module_root_dir)/../lib", "-lvgcut", "-lvgdag", "-Wl,-rpath,<(module_root_dir)/../lib"],
    "cflags_cc": ["-std=c++17"]
  }]
}
