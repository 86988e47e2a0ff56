//! Link against libnqpu_hip.so.
//!
//! The library is built by `python -m nqpu_amd.build` (hipcc, gfx950) into
//! `nqpu_amd/lib/` of this repository.  `NQPU_HIP_LIB_DIR` overrides the
//! directory (for an installed copy).  The run-time search path is set to the
//! same directory so test binaries find the library without LD_LIBRARY_PATH.
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=NQPU_HIP_LIB_DIR");
    let dir = match env::var("NQPU_HIP_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => {
            let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR"));
            // rust/nqpu-hip-sys -> <repo>/nqpu_amd/lib
            manifest.join("..").join("..").join("nqpu_amd").join("lib")
        }
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=nqpu_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
    // ROCm's runtime libraries are found through the library's own DT_NEEDED
    // entries (libamdhip64, libhiprtc) and the system's ROCm installation.
}
