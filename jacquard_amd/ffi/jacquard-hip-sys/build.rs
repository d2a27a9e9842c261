//! Links libjacquard_hip.so (built by `python -m jacquard_amd.build` into
//! jacquard_amd/lib/) and the HIP runtime.  This replaces the HIP branch of
//! the reference's build script (build.rs:74-107), which compiles
//! csrc/kernel_v1.hip.cpp into the static library `gemhip`: here the device
//! code is prebuilt, so the crate only links.
use std::env;
use std::path::PathBuf;

fn main() {
    let lib_dir = env::var("JACQUARD_HIP_LIB_DIR").map(PathBuf::from).unwrap_or_else(|_| {
        // default: the library built in place, two levels above this crate
        PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../lib")
    });
    let rocm = env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string());
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-search=native={}/lib", rocm);
    println!("cargo:rustc-link-lib=dylib=jacquard_hip");
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    // find both at run time without LD_LIBRARY_PATH
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}/lib", rocm);
    println!("cargo:rerun-if-env-changed=JACQUARD_HIP_LIB_DIR");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");
}
