// Link librcgi.so from the directory RC_GI_LIB_DIR names (the repository's
// radiancecascade_amd/ after `python __graft_entry__.py`), with the HIP
// runtime it needs from ROCm's lib directory.
fn main() {
    let lib_dir = std::env::var("RC_GI_LIB_DIR").unwrap_or_else(|_| "../../../radiancecascade_amd".to_string());
    let rocm = std::env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string());
    println!("cargo:rustc-link-search=native={lib_dir}");
    println!("cargo:rustc-link-search=native={rocm}/lib");
    println!("cargo:rustc-link-lib=dylib=rcgi");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{lib_dir}");
    println!("cargo:rerun-if-env-changed=RC_GI_LIB_DIR");
}
