// Link against libpyro_hip.so.  PYRO_HIP_LIB_DIR names the folder that holds
// it (default: pyro_amd/lib of this repository, three levels up).
fn main() {
    let dir = std::env::var("PYRO_HIP_LIB_DIR").unwrap_or_else(|_| {
        let here = std::path::PathBuf::from(std::env::var("CARGO_MANIFEST_DIR").unwrap());
        here.join("../../../pyro_amd/lib").to_string_lossy().into_owned()
    });
    println!("cargo:rerun-if-env-changed=PYRO_HIP_LIB_DIR");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=pyro_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
}
