//! Links libvistio_hip.so, built by `make -C vistio_amd/csrc` (hipcc,
//! --offload-arch=gfx950). VISTIO_HIP_LIB_DIR names the folder that holds it
//! (the vistio_amd/ folder of the HIP build); the run-time loader finds it
//! through the rpath set here or LD_LIBRARY_PATH.
fn main() {
    println!("cargo:rerun-if-env-changed=VISTIO_HIP_LIB_DIR");
    if let Ok(dir) = std::env::var("VISTIO_HIP_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    }
    println!("cargo:rustc-link-lib=dylib=vistio_hip");
}
