// Links libarp.so: ARP_LIB_DIR names the folder that holds it (by default the library built
// in-tree by ar-physics_amd/build.py, relative to this crate), and the HIP runtime it needs.
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = env::var("ARP_LIB_DIR").map(PathBuf::from).unwrap_or_else(|_| {
        PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../ar-physics_amd/arp")
    });
    let rocm = env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string());
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-search=native={}/lib", rocm);
    println!("cargo:rustc-link-lib=dylib=arp");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
    println!("cargo:rerun-if-env-changed=ARP_LIB_DIR");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");
}
