//! Builds libholomap_gpu.so from this repository's HIP sources with hipcc
//! (the flags of holomap_amd/build.py) and links it.
//!
//! HOLOMAP_GPU_SRC overrides the source directory (default: the
//! holomap_amd/csrc directory two levels above this crate); HIPCC the
//! compiler (default /opt/rocm/bin/hipcc).
use std::{env, path::PathBuf, process::Command};

const SOURCES: [&str; 9] = [
    "hm_prep", "hm_gram", "hm_knn", "hm_calib", "hm_sort", "hm_fuzzy", "hm_graph", "hm_runtime",
    "hm_serve",
];

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let src = env::var("HOLOMAP_GPU_SRC")
        .map(PathBuf::from)
        .unwrap_or_else(|_| manifest.join("../../holomap_amd/csrc"));
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let hipcc = env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".into());
    // numerics: no contraction (the reference's f64 folds and f32
    // calibration), correctly rounded f32 division / sqrt, no fast-math;
    // no packed-FP32 instructions
    let flags = [
        "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    ];
    let mut objs = Vec::new();
    for f in SOURCES {
        let file = src.join(format!("{f}.hip"));
        let obj = out.join(format!("{f}.o"));
        let st = Command::new(&hipcc).args(flags).arg("-c").arg(&file).arg("-o").arg(&obj)
            .status().expect("run hipcc");
        assert!(st.success(), "hipcc failed on {}", file.display());
        println!("cargo:rerun-if-changed={}", file.display());
        objs.push(obj);
    }
    for h in ["hm_internal.hpp", "hm_scan.hpp"] {
        println!("cargo:rerun-if-changed={}", src.join(h).display());
    }
    println!("cargo:rerun-if-changed={}", src.join("../../include/holomap_gpu.h").display());
    let lib = out.join("libholomap_gpu.so");
    let st = Command::new(&hipcc).args(["--offload-arch=gfx950", "-shared", "-fPIC", "-o"])
        .arg(&lib).args(&objs).status().expect("run hipcc");
    assert!(st.success(), "link failed");
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=holomap_gpu");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
    println!("cargo:rustc-link-lib=dylib=amdhip64");
}
