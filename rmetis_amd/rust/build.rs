//! Builds librmetis_hip.so from ../csrc with hipcc for gfx950 and links it.
//! HIPCC (default /opt/rocm/bin/hipcc) and RMETIS_HIP_ARCH (default gfx950)
//! may be set in the environment.
use std::env;
use std::path::PathBuf;
use std::process::Command;

// The library's translation units, in the order rmetis_amd/build.py uses.
const SOURCES: &[&str] = &[
    "scan.hip", "reduce.hip", "refine.hip", "admit.hip", "coarsen.hip", "fm.hip", "fm2.hip", "grow.hip",
    "split.hip", "tuning.cpp", "driver.cpp", "batch.cpp", "recursive.cpp", "graphio.cpp", "serve.cpp", "capi.cpp",
];

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let csrc = manifest.join("..").join("csrc");
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let hipcc = env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".to_string());
    let arch = env::var("RMETIS_HIP_ARCH").unwrap_or_else(|_| "gfx950".to_string());
    let lib = out.join("librmetis_hip.so");
    let mut cmd = Command::new(&hipcc);
    cmd.arg(format!("--offload-arch={}", arch))
        .args(["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off"])
        // packed FP32 instructions are kept out of every kernel
        .args(["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"])
        .args(["-x", "hip"]);
    for s in SOURCES {
        cmd.arg(csrc.join(s));
    }
    cmd.arg("-o").arg(&lib);
    let status = cmd.status().unwrap_or_else(|e| panic!("cannot run {}: {}", hipcc, e));
    assert!(status.success(), "hipcc failed building librmetis_hip.so");
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=rmetis_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
    println!("cargo:rerun-if-changed={}", csrc.display());
    println!("cargo:rerun-if-changed={}", manifest.join("..").join("..").join("include").display());
    println!("cargo:rerun-if-env-changed=HIPCC");
    println!("cargo:rerun-if-env-changed=RMETIS_HIP_ARCH");
}
