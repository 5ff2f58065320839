// build.rs -- link libaemh.so, the way ae_cpp_kernel's build.rs
// (ae_cpp_kernel/build.rs:11-19) compiles and links its C++ kernel.
//
// The HIP library is built by its own Makefile with hipcc for gfx950 (the cc
// crate has no HIP target), so this script runs that Makefile unless
// AEMH_LIB_DIR names a directory that already holds libaemh.so.
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    println!("cargo:rerun-if-env-changed=AEMH_LIB_DIR");
    let lib_dir = match env::var("AEMH_LIB_DIR") {
        Ok(dir) => PathBuf::from(dir),
        Err(_) => {
            // this crate sits in ae-engine_amd/ae_mpm_hip; the sources in
            // ae-engine_amd/csrc, the library goes to ae-engine_amd/aemh
            let here = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
            let csrc = here.join("../csrc");
            let status = Command::new("make")
                .arg("-C")
                .arg(&csrc)
                .status()
                .expect("make (hipcc) for libaemh.so");
            assert!(status.success(), "building libaemh.so failed");
            for f in ["aemh_api.hip", "kernels.inc.h", "aemh_physics.h", "../../include/aemh.h"] {
                println!("cargo:rerun-if-changed={}", csrc.join(f).display());
            }
            here.join("../aemh")
        }
    };
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=aemh");
    // find the library at run time without LD_LIBRARY_PATH
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
}
