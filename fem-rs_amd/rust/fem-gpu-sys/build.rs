//! Builds libfemgpu.so with the library's own Makefile (hipcc
//! --offload-arch=gfx950) and links it.  FEMGPU_LIB_DIR points at a prebuilt
//! library instead; FEMGPU_SKIP_BUILD=1 links without building.
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    // fem-rs_amd/rust/fem-gpu-sys -> fem-rs_amd
    let pkg = manifest.join("..").join("..");
    let lib_dir = match env::var("FEMGPU_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => {
            if env::var("FEMGPU_SKIP_BUILD").map(|v| v != "1").unwrap_or(true) {
                let status = Command::new("make")
                    .arg("-C")
                    .arg(&pkg)
                    .status()
                    .expect("running make for libfemgpu (needs hipcc from ROCm)");
                assert!(status.success(), "building libfemgpu failed");
            }
            pkg.join("femgpu")
        }
    };
    println!("cargo:rerun-if-env-changed=FEMGPU_LIB_DIR");
    println!("cargo:rerun-if-env-changed=FEMGPU_SKIP_BUILD");
    println!("cargo:rerun-if-changed={}", pkg.join("csrc").display());
    println!("cargo:rerun-if-changed={}", pkg.join("..").join("include").join("femgpu.h").display());
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=femgpu");
    // the library is found at run time next to where it was built
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
}
