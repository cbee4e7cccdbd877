"""pyemma.coordinates-compatible entry points of the GPU discretization path."""
from .api import assign_to_centers, cluster_kmeans, covariance_lagged, pca, tica, tica_nystroem, vamp
from .clustering import AssignCenters, KmeansClustering
from .covariance import LaggedCovariance
from .featurizer import Featurizer, featurizer
from .nystroem import NystroemTICA
from .pca import PCA
from .source import FeatureSource, source
from .tica import TICA
from .vamp import VAMP

__all__ = ["tica", "covariance_lagged", "cluster_kmeans", "assign_to_centers", "vamp", "pca", "TICA",
           "LaggedCovariance", "KmeansClustering", "AssignCenters", "VAMP", "PCA", "Featurizer", "featurizer", "tica_nystroem",
           "NystroemTICA", "FeatureSource", "source"]
