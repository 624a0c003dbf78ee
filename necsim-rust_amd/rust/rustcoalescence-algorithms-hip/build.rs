// Links the prebuilt C-ABI library (necsim-rust_amd/libnecsim_hip.so).
fn main() {
    let dir = std::env::var("NECSIM_HIP_LIB_DIR").unwrap_or_else(|_| "../..".into());
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=necsim_hip");
    println!("cargo:rerun-if-env-changed=NECSIM_HIP_LIB_DIR");
}
