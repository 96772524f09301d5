//! Build step for `libgnomon_hip.so`, called from gnomon's build script.
//!
//! NOT BUILT IN THIS IMAGE: there is no cargo here.  In gnomon, `Cargo.toml:15`
//! names `shared/build.rs` as the build script; add `mod hip_build;` (this
//! file, placed beside it) and call `hip_build::build_gnomon_hip()` from its
//! `main()`.  The flags and the source list are those of
//! `gnomon_amd/build.py` (`HIPCC_FLAGS`, `LINK_FLAGS`, `SOURCE_NAMES`), which
//! builds the same library in this repository; `tests/test_abi.py` checks
//! that the lists agree.  Code objects are compiled ahead of time for gfx950;
//! nothing is compiled at run time.

use std::path::{Path, PathBuf};
use std::process::Command;

/// Every translation unit of the library, in link order.
const SOURCES: [&str; 7] = [
    "gnm_kernels.hip",
    "gnm_dense.hip",
    "gnm_i8.hip",
    "gnm_bf16.hip",
    "gnm_sparse.hip",
    "gnm_plan.hip",
    "gnm_api.hip",
];

const COMPILE_FLAGS: [&str; 11] = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    // hipcc -O3 may otherwise form packed FP32 instructions from scalar code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-c",
];

const LINK_FLAGS: [&str; 4] = ["--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined"];

/// `GNOMON_HIP_SRC` points at this repository's root (the directory holding
/// `include/` and `gnomon_amd/csrc/`); defaults to `hip/` inside the crate.
pub fn build_gnomon_hip() {
    let out = PathBuf::from(std::env::var("OUT_DIR").expect("OUT_DIR"));
    let root = PathBuf::from(std::env::var("GNOMON_HIP_SRC").unwrap_or_else(|_| "hip".to_string()));
    let csrc = root.join("gnomon_amd").join("csrc");
    let include = root.join("include");
    let hipcc = std::env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".to_string());

    let mut objects = Vec::with_capacity(SOURCES.len());
    for name in SOURCES {
        let src = csrc.join(name);
        let obj = out.join(format!("{name}.o"));
        run(Command::new(&hipcc).args(COMPILE_FLAGS).arg(format!("-I{}", include.display()))
            .arg(&src).arg("-o").arg(&obj), name);
        objects.push(obj);
    }
    let lib = out.join("libgnomon_hip.so");
    run(Command::new(&hipcc).args(LINK_FLAGS).args(&objects).arg("-o").arg(&lib), "link");

    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=gnomon_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
    rerun_if_changed(&csrc);
    rerun_if_changed(&include);
}

fn run(cmd: &mut Command, what: &str) {
    let status = cmd.status().unwrap_or_else(|e| panic!("hipcc ({what}) could not start: {e}"));
    assert!(status.success(), "hipcc ({what}) failed with {status}");
}

fn rerun_if_changed(dir: &Path) {
    println!("cargo:rerun-if-changed={}", dir.display());
    println!("cargo:rerun-if-env-changed=GNOMON_HIP_SRC");
    println!("cargo:rerun-if-env-changed=HIPCC");
}
