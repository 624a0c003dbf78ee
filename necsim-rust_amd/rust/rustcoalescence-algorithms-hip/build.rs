// Links the prebuilt C-ABI library libnecsim_hip.so (built with
// necsim-rust_amd/build.py); NECSIM_HIP_LIB_DIR names its directory.
fn main() {
    println!("cargo:rerun-if-env-changed=NECSIM_HIP_LIB_DIR");
    let dir = std::env::var("NECSIM_HIP_LIB_DIR").unwrap_or_else(|_| {
        panic!("set NECSIM_HIP_LIB_DIR to the directory that holds libnecsim_hip.so (necsim-rust_amd/)")
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=necsim_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
}
